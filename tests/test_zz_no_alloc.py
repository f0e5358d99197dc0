"""The allocation-free symmetric-Dirichlet energy (rxm_symdir_energy_noalloc)
against the allocating entry point (rxm_symdir_energy), both through the
C-ABI: the same kernels, so totals and per-face energies must be bit-equal.
Two calls in a row, the second on a smaller mesh, so the second call runs
over what the first left in the library's fixed scratch."""
import numpy as np
import pytest

from rxmesh_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _setup(V, F, patch_size, seed):
    import torch
    import rxmesh_amd as rx
    m = rx.Mesh(F, patch_size=patch_size)
    l2g = m.array("v_l2g")
    lin = lambda X: torch.from_numpy(np.ascontiguousarray(X[l2g].T.astype(np.float32))).cuda()  # noqa: E731
    uv = mg.uv_init(V, F, seed=seed).astype(np.float32)
    Mr = m.symdir_rest(lin(V.astype(np.float32)))
    return m, Mr, lin(uv)


def _both(m, Mr, uv):
    import torch
    t_ref, e_ref = m.symdir_energy(Mr, uv)
    t_na, e_na = m.symdir_energy(Mr, uv, noalloc=True)
    torch.cuda.synchronize()
    return (float(t_ref), e_ref.cpu().numpy()), (float(t_na), e_na.cpu().numpy())


def test_noalloc_energy_equals_allocating_entry_point():
    import torch
    big = _setup(*mg.config_mesh("c3"), patch_size=128, seed=0)   # ~100k faces, ~800 patches
    small = _setup(*mg.disk(48, seed=3), patch_size=64, seed=1)   # a few thousand faces
    assert big[0].P > small[0].P
    print("noalloc: big mesh P=%d F=%d, small mesh P=%d F=%d" % (big[0].P, big[0].nF, small[0].P, small[0].nF),
          flush=True)
    for m, Mr, uv in (big, small):
        (t_ref, e_ref), (t_na, e_na) = _both(m, Mr, uv)
        assert np.isfinite(t_ref) and t_ref > 0
        assert t_na == t_ref, (t_na, t_ref)
        assert np.array_equal(e_na, e_ref)
    # the small mesh again after the big one overwrote the scratch with more partials
    m, Mr, uv = big
    m.symdir_energy(Mr, uv, noalloc=True)
    m, Mr, uv = small
    (t_ref, _), (t_na, _) = _both(m, Mr, uv)
    assert t_na == t_ref
    torch.cuda.synchronize()
    for m, _, _ in (big, small):
        m.close()
