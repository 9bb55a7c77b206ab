"""bfm_dev_get_velocities_noalloc (no device allocation, no context) against
the entry points it stands for, bfm_dev_import_neighbors +
bfm_dev_get_velocities: bit-identical velocities on two calls in a row, the
second with a smaller input (the library's fixed working array keeps the
first call's contents), and the oracle's getVelocities within 1e-10."""
import numpy as np
import pytest

from gpu_common import LDIV, R, THR, rel_err

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asfortranarray(a).T)).cuda()


def _reference_call(cells, nb, p, a2d=None, noise=None):
    import torch

    import bfm
    n, m = cells.shape[0], nb.shape[0]
    ctx = bfm.Context(n, max_pairs=max(m, 1024), params=p)
    d_cells, d_nb = _dev(cells), _dev(nb)
    d_vel = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    d_a2d = torch.from_numpy(a2d).cuda() if a2d is not None else None
    d_noise = _dev(noise) if noise is not None else None
    torch.cuda.synchronize()
    ctx.dev_import_neighbors(d_nb.data_ptr(), m, m, n)
    ctx.dev_get_velocities(d_cells.data_ptr(), n, n, d_vel.data_ptr(),
                           d_a2d.data_ptr() if d_a2d is not None else 0,
                           d_noise.data_ptr() if d_noise is not None else 0)
    ctx.sync()
    torch.cuda.synchronize()
    return d_vel.cpu().numpy().T.copy()


def _noalloc_call(cells, nb, p, a2d=None, noise=None):
    import torch

    import bfm
    n, m = cells.shape[0], nb.shape[0]
    d_cells, d_nb = _dev(cells), _dev(nb)
    d_vel = torch.full((6, n), float("nan"), dtype=torch.float64, device="cuda")
    d_a2d = torch.from_numpy(a2d).cuda() if a2d is not None else None
    d_noise = _dev(noise) if noise is not None else None
    torch.cuda.synchronize()
    bfm.core.dev_get_velocities_noalloc(p, d_cells.data_ptr(), n, n, d_nb.data_ptr(), m, m,
                                        d_vel.data_ptr(),
                                        d_a2d.data_ptr() if d_a2d is not None else 0,
                                        d_noise.data_ptr() if d_noise is not None else 0)
    torch.cuda.synchronize()
    return d_vel.cpu().numpy().T.copy()


def test_noalloc_matches_existing_entry_twice(orc):
    import bfm
    p = bfm.core.make_params()
    big, _ = bfm.make_dome(40000, seed=3)
    small, _ = bfm.make_dome(6000, seed=9)
    for cells in (big, small):
        n = cells.shape[0]
        nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
        print(f"noalloc n={n} m={nb.shape[0]}", flush=True)
        a2d = (cells[:, 6] < p.nz_threshold).astype(np.int32)
        want = _reference_call(cells, nb, p)
        got = _noalloc_call(cells, nb, p)
        assert np.array_equal(got, want)
        v_ref = orc.get_velocities(cells, nb, a2d, None, p)
        assert rel_err(got, v_ref) <= 1e-10


def test_noalloc_noise_assume2d_and_hertz(orc):
    import bfm
    cells, _ = bfm.make_dome(3000, seed=5)
    n = cells.shape[0]
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    rng = np.random.default_rng(1)
    noise = np.asfortranarray(rng.normal(scale=10.0, size=(n, 6)))
    a2d = (rng.uniform(size=n) < 0.2).astype(np.int32)
    for p in (bfm.core.make_params(), bfm.core.make_params(adhesion=False)):
        print(f"noalloc noise adhesion={p.adhesion_mode}", flush=True)
        want = _reference_call(cells, nb, p, a2d, noise)
        got = _noalloc_call(cells, nb, p, a2d, noise)
        assert np.array_equal(got, want)
        assert rel_err(got, orc.get_velocities(cells, nb, a2d, noise, p)) <= 1e-10


def test_noalloc_rejects_sizes_and_modes():
    import bfm
    from bfm.core import BFMError, BFM_ERR_CAPACITY, BFM_ERR_UNSUPPORTED
    p = bfm.core.make_params()
    with pytest.raises(BFMError) as e:
        bfm.core.dev_get_velocities_noalloc(p, 1, 65537, 65537, 1, 10, 10, 1)
    assert e.value.status == BFM_ERR_CAPACITY
    with pytest.raises(BFMError) as e:
        bfm.core.dev_get_velocities_noalloc(p, 1, 100, 100, 1, 262145, 262145, 1)
    assert e.value.status == BFM_ERR_CAPACITY
    for q in (bfm.core.make_params(jkr_mode=bfm.core.JKR_TABLE_1D), bfm.core.make_params(friction=True)):
        with pytest.raises(BFMError) as e:
            bfm.core.dev_get_velocities_noalloc(q, 1, 100, 100, 1, 10, 10, 1)
        assert e.value.status == BFM_ERR_UNSUPPORTED
