"""The device edge build (rxm_build_edges_device) against the oracle's
restatement of the reference's edge numbering (first appearance over the
half-edges, first-seen orientation; itself pinned to the reference's patch
archives and a brute force in test_oracle.py), bit for bit, and against the
product's host build."""
import numpy as np
import pytest

from conftest import obj
from oracle.oracle import OracleMesh
from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _faces(name):
    if name.endswith(".obj"):
        return obj(name)[1]
    if name == "ico_perm":
        return mg.permute_mesh(*mg.icosphere(4, seed=2, sigma=0.005), seed=3)[1]
    if name == "torus":
        return mg.torus(40, 24)[1]
    if name == "disk":
        return mg.disk(30, seed=5)[1]
    if name == "c1":
        return mg.config_mesh("c1")[1]
    if name == "flipped":
        F = mg.config_mesh("c1")[1].copy()
        F[7] = F[7][::-1]
        return F
    if name == "ico_odd":  # an odd face count: 3F half-edges is odd
        return mg.icosphere(4, seed=2, sigma=0.005)[1][:-1]
    if name == "fin":  # an edge with three faces
        return np.array([[0, 1, 2], [1, 0, 3], [0, 1, 4], [2, 1, 5]], np.uint32)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["plane.obj", "cube.obj", "sphere3.obj", "bunnyhead.obj", "ico_perm", "torus",
                                  "disk", "c1", "flipped", "fin", "ico_odd"])
def test_device_edges_equal_host_numbering(name):
    import torch
    import rxmesh_amd as rx
    F = np.ascontiguousarray(_faces(name), dtype=np.uint32)
    nV = int(F.max()) + 1
    print("case", name, len(F), flush=True)
    ev, fe = rx.build_edges_device(torch.from_numpy(F.astype(np.int32)).cuda(), nV)
    torch.cuda.synchronize()
    o = OracleMesh(F)
    ref = o.ev().astype(np.int64)
    assert np.array_equal(ev.cpu().numpy().astype(np.int64), ref)
    ids = {tuple(sorted(e)): i for i, e in enumerate(ref.tolist())}
    exp = [ids[tuple(sorted((f[i], f[(i + 1) % 3])))] for f in F.tolist() for i in range(3)]
    assert np.array_equal(fe.cpu().numpy(), np.array(exp))
    if name != "fin":
        m = rx.Mesh(F, host_only=True)
        assert np.array_equal(m.array("ev").reshape(-1, 2).astype(np.int64), ref)


def test_device_edges_reject_bad_faces():
    import torch
    import rxmesh_amd as rx
    F = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    with pytest.raises(rx.RXMError) as e:
        rx.build_edges_device(torch.from_numpy(F).cuda(), 3)  # vertex 3 out of range
    assert e.value.code == 1
    F = np.array([[0, 1, 2], [1, 1, 2]], np.int32)
    with pytest.raises(rx.RXMError) as e:
        rx.build_edges_device(torch.from_numpy(F).cuda(), 3)  # degenerate face
    assert e.value.code == 1


@pytest.mark.parametrize("name", ["plane.obj", "cube.obj", "sphere3.obj", "bunnyhead.obj", "ico_perm", "torus",
                                  "disk", "c1", "flipped", "fin", "ico_odd"])
def test_mesh_create_device_connectivity_equals_host_build(name):
    """rxm_mesh_create builds the edges, EV and the VF / EF CSRs on the
    device; every array the build exports, the mesh statistics and the
    patch layout equal the host-only build's."""
    import rxmesh_amd as rx
    F = np.ascontiguousarray(_faces(name), dtype=np.uint32)
    print("case", name, len(F), flush=True)
    d = rx.Mesh(F, patch_size=64)
    h = rx.Mesh(F, patch_size=64, host_only=True)
    for key in ("num_vertices", "num_edges", "num_faces", "num_patches", "is_closed", "is_edge_manifold",
                "is_consistently_oriented", "max_valence", "max_edge_incident_faces", "max_face_adjacent_faces"):
        if key in h.info:
            assert d.info[key] == h.info[key], key
    for arr in rx.ARR:
        if arr.startswith("vv2"):
            continue
        assert np.array_equal(d.array(arr), h.array(arr)), arr
    d.close()
    h.close()


def test_concurrent_device_builds_equal_serial_builds():
    """Meshes built on the device path from several threads at once (each
    build has its own stream and buffers; the C-ABI releases the GIL; host
    side on the calling thread only, serial_build) and a batch whose meshes
    are built in parallel export the same arrays as the same meshes built one
    by one with every host thread."""
    from concurrent.futures import ThreadPoolExecutor
    import rxmesh_amd as rx
    meshes = [mg.icosphere(4, seed=k, sigma=0.01)[1] for k in range(3)] + [mg.grid(40 + 7 * k, seed=k)[1]
                                                                          for k in range(3)]
    meshes += [mg.permute_mesh(*mg.config_mesh("c1"), seed=5)[1], _faces("torus"), _faces("disk")]
    meshes = [np.ascontiguousarray(f, dtype=np.uint32) for f in meshes]
    serial = [rx.Mesh(f, patch_size=256) for f in meshes]
    with ThreadPoolExecutor(max_workers=4) as pool:
        conc = list(pool.map(lambda f: rx.Mesh(f, patch_size=256, serial_build=True), meshes))
    names = [a for a in rx.ARR if not a.startswith("vv2")]
    for a, b in zip(serial, conc):
        for n in names:
            assert np.array_equal(a.array(n), b.array(n)), n
    batch = rx.Mesh(None, batch=meshes, patch_size=256)
    union = rx.Mesh(None, parts=serial, host_only=True)
    for n in names:
        assert np.array_equal(batch.array(n), union.array(n)), n
