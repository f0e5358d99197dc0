"""Seeded synthetic meshes for the benchmark workloads and tests.

Random numbers come from a counter-based splitmix64 stream (Steele, Lea,
Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014), so a
(seed, stream, index) triple always yields the same value. Normal deviates
use the Box-Muller transform.
"""
import math

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = x.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        x += np.uint64(0x9E3779B97F4A7C15)
        z = x
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, stream, n):
    """n uniforms in [0, 1) for (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * np.uint64(0x100000001B3) + np.uint64(stream) * np.uint64(0x9E3779B97F4A7C15)
        idx = np.arange(n, dtype=np.uint64) + base
    z = _splitmix64(_splitmix64(idx))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, stream, n):
    """n standard normals (Box-Muller over two uniform streams)."""
    u1 = uniform(seed, 2 * stream + 1, n)
    u2 = uniform(seed, 2 * stream + 2, n)
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def permutation(seed, stream, n):
    """Seeded random permutation of range(n)."""
    return np.argsort(uniform(seed, stream, n), kind="stable")


# ---------------------------------------------------------------------------
def grid(n, seed=0, jitter=0.25, wave=True):
    """n x n vertices on [0,1]^2, cell (i,j) split along a-d when (i+j) is
    even, else along b-c; CCW faces; interior vertices jittered by
    U(-jitter*h, jitter*h) in x, y; z = 0.05 sin(6x) cos(6y)."""
    h = 1.0 / (n - 1)
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    x = (i * h).astype(np.float64).ravel()
    y = (j * h).astype(np.float64).ravel()
    interior = ((i > 0) & (i < n - 1) & (j > 0) & (j < n - 1)).ravel()
    if jitter:
        nv = n * n
        ux = uniform(seed, 1, nv) * 2.0 - 1.0
        uy = uniform(seed, 2, nv) * 2.0 - 1.0
        x = np.where(interior, x + jitter * h * ux, x)
        y = np.where(interior, y + jitter * h * uy, y)
    z = 0.05 * np.sin(6.0 * x) * np.cos(6.0 * y) if wave else np.zeros_like(x)
    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="xy")
    ci = ci.ravel()
    cj = cj.ravel()
    a = cj * n + ci
    b = a + 1
    c = a + n
    d = c + 1
    even = ((ci + cj) % 2) == 0
    f0 = np.where(even[:, None], np.stack([a, b, d], 1), np.stack([a, b, c], 1))
    f1 = np.where(even[:, None], np.stack([a, d, c], 1), np.stack([b, d, c], 1))
    fv = np.empty((2 * len(a), 3), np.int64)
    fv[0::2] = f0
    fv[1::2] = f1
    return np.stack([x, y, z], 1), fv.astype(np.uint32)


def icosphere(level, seed=0, sigma=0.0):
    """Icosahedron subdivided `level` times by midpoint split, projected to
    the unit sphere; radial noise r = 1 + N(0, sigma)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    for _ in range(level):
        nV = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.maximum(e[:, 0], e[:, 1]) * nV + np.minimum(e[:, 0], e[:, 1])
        uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(len(order))
        mid_id = nV + rank[inv]
        ek = e[first[order]]
        mp = v[ek[:, 0]] + v[ek[:, 1]]
        mp /= np.linalg.norm(mp, axis=1, keepdims=True)
        v = np.concatenate([v, mp])
        F = len(f)
        mab, mbc, mca = mid_id[:F], mid_id[F:2 * F], mid_id[2 * F:]
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.stack([np.stack([a, mab, mca], 1), np.stack([b, mbc, mab], 1),
                      np.stack([c, mca, mbc], 1), np.stack([mab, mbc, mca], 1)], 1).reshape(-1, 3)
    if sigma:
        v = v * (1.0 + sigma * normal(seed, 3, len(v)))[:, None]
    return v, f.astype(np.uint32)


def torus(nu, nw, R=1.0, r=0.3):
    """nu x nw torus grid; quads split by the (i+j) parity rule of grid()."""
    i, j = np.meshgrid(np.arange(nu), np.arange(nw), indexing="ij")
    th = 2.0 * math.pi * i.ravel() / nu
    ph = 2.0 * math.pi * j.ravel() / nw
    pos = np.stack([(R + r * np.cos(ph)) * np.cos(th), (R + r * np.cos(ph)) * np.sin(th), r * np.sin(ph)], 1)
    vid = lambda a, b: (a % nu) * nw + (b % nw)  # noqa: E731
    ci, cj = i.ravel(), j.ravel()
    a = vid(ci, cj)
    b = vid(ci + 1, cj)
    c = vid(ci, cj + 1)
    d = vid(ci + 1, cj + 1)
    even = ((ci + cj) % 2) == 0
    f0 = np.where(even[:, None], np.stack([a, b, d], 1), np.stack([a, b, c], 1))
    f1 = np.where(even[:, None], np.stack([a, d, c], 1), np.stack([b, d, c], 1))
    fv = np.empty((2 * len(a), 3), np.int64)
    fv[0::2] = f0
    fv[1::2] = f1
    return pos, fv.astype(np.uint32)


def permute_mesh(pos, fv, seed, vstream=11, fstream=12):
    """Randomly relabel vertices and faces (destroys input locality)."""
    nv, nf = len(pos), len(fv)
    pv = permutation(seed, vstream, nv)       # new id -> old id
    inv = np.empty(nv, np.int64)
    inv[pv] = np.arange(nv)
    pf = permutation(seed, fstream, nf)
    return pos[pv], inv[fv[pf]].astype(np.uint32)


def disk(n=256, seed=0):
    """Jittered n x n grid clipped to the disk |(x,y) - (.5,.5)| <= .5, with
    compacted ids."""
    pos, fv = grid(n, seed=seed)
    inside = np.hypot(pos[:, 0] - 0.5, pos[:, 1] - 0.5) <= 0.5
    keep = inside[fv].all(axis=1)
    fv = fv[keep]
    used = np.zeros(len(pos), bool)
    used[fv.ravel()] = True
    newid = -np.ones(len(pos), np.int64)
    newid[used] = np.arange(used.sum())
    return pos[used], newid[fv].astype(np.uint32)


def uv_init(pos, fv, seed=0, sigma=0.002, tries=16):
    """UV = rest (x, y) + N(0, sigma), re-drawn for vertices of faces whose
    orientation would not stay positive."""
    base = pos[:, :2].astype(np.float64)
    uv = base + sigma * np.stack([normal(seed, 21, len(pos)), normal(seed, 22, len(pos))], 1)
    for t in range(tries):
        a, b, c = uv[fv[:, 0]], uv[fv[:, 1]], uv[fv[:, 2]]
        det = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (c[:, 0] - a[:, 0]) * (b[:, 1] - a[:, 1])
        bad = det <= 0
        if not bad.any():
            break
        vs = np.unique(fv[bad].ravel())
        redraw = sigma * np.stack([normal(seed, 23 + 2 * t, len(pos)), normal(seed, 24 + 2 * t, len(pos))], 1)
        uv[vs] = base[vs] + redraw[vs] * (0.5 ** (t + 1))
    return uv


def batch_specs(count=256, seed=0, fmin=2e3, fmax=1.3e6):
    """Mesh list of the batched workload: face counts log-uniform in
    [fmin, fmax]; even entries icospheres (level 3..8 nearest in log F), odd
    entries jittered n x n grids (n in [32, 800])."""
    u = uniform(seed, 31, count)
    F = np.exp(np.log(fmin) + u * (np.log(fmax) - np.log(fmin)))
    out = []
    for k in range(count):
        if k % 2 == 0:
            lev = int(np.clip(np.round(np.log(F[k] / 20.0) / np.log(4.0)), 3, 8))
            out.append(("icosphere", lev))
        else:
            n = int(np.clip(round(math.sqrt(F[k] / 2.0)) + 1, 32, 800))
            out.append(("grid", n))
    return out


def make_batch_mesh(spec, seed):
    kind, arg = spec
    if kind == "icosphere":
        return icosphere(arg, seed=seed, sigma=0.005)
    return grid(arg, seed=seed)


def config_mesh(name, seed=0):
    """Named workload meshes: c0 grid128, c1 icosphere L6 permuted, c2s
    icosphere L9, c2t torus 2048x1024, c3 disk."""
    if name == "c0":
        return grid(128, seed=seed)
    if name == "c1":
        pos, fv = icosphere(6, seed=seed, sigma=0.01)
        return permute_mesh(pos, fv, seed)
    if name == "c2s":
        return icosphere(9, seed=seed, sigma=0.005)
    if name == "c2s_perm":
        pos, fv = icosphere(9, seed=seed, sigma=0.005)
        return permute_mesh(pos, fv, seed)
    if name == "c2t":
        return torus(2048, 1024)
    if name == "c3":
        return disk(256, seed=seed)
    raise KeyError(name)


def read_obj(path):
    """Minimal OBJ reader: 'v' positions and triangular 'f' faces (1-based,
    'a/b/c' tokens allowed)."""
    V, F = [], []
    with open(path) as fh:
        for line in fh:
            s = line.split()
            if not s:
                continue
            if s[0] == "v":
                V.append([float(x) for x in s[1:4]])
            elif s[0] == "f":
                idx = [int(t.split("/")[0]) for t in s[1:]]
                idx = [k - 1 if k > 0 else len(V) + k for k in idx]
                if len(idx) != 3:
                    raise ValueError("non-triangular face in %s" % path)
                F.append(idx)
    return np.array(V, np.float64), np.array(F, np.uint32)
