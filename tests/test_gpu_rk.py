"""GPU RK step (stepRungeKuttaAdaptive, mechanics.hpp:2406-2612) and the
device-resident stage against the oracle. Bar: 1e-10 normwise per column for
positions, orientations and stored velocities; embedded errors within 1e-10 of
the step's displacement scale."""
import numpy as np
import pytest

from gpu_common import LDIV, R, THR, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dome():
    import bfm
    from oracle import oracle as orc
    cells, _ = bfm.make_dome(8000, seed=31)
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    return cells, nb


@pytest.mark.parametrize("name", ["heun_euler", "bogacki_shampine", "rkf45", "dormand_prince"])
def test_step_rk_tableaux(orc, dome, name):
    import bfm
    cells, nb = dome
    p = bfm.core.make_params()
    A, b, bs, _ = bfm.tableau(name)
    c = bfm.Context(8000, params=p)
    rng = np.random.default_rng(32)
    noise = np.asfortranarray(rng.uniform(-1, 1, (8000, 6)))
    for nz in (None, noise):
        new, err = c.step_rk(A, b, bs, 2e-5, cells, nb, nz)
        new_ref, err_ref = orc.step_rk(A, b, bs, 2e-5, cells, nb, p, nz)
        assert rel_err(new[:, 1:13], new_ref[:, 1:13]) <= 1e-10
        assert np.array_equal(new[:, [0] + list(range(13, 22))],
                              new_ref[:, [0] + list(range(13, 22))])
        disp = np.max(np.abs(new_ref[:, 1:7] - cells[:, 1:7]), axis=0)
        assert np.all(np.max(np.abs(err - err_ref), axis=0) <= 1e-10 * disp)


def test_device_step_and_scaled_error(orc, dome):
    import torch

    import bfm
    cells, nb = dome
    p = bfm.core.make_params()
    A, b, bs, _ = bfm.tableau("dormand_prince")
    n = cells.shape[0]
    c = bfm.Context(n, params=p)
    d_cells = torch.from_numpy(np.ascontiguousarray(cells.T)).cuda()
    d_new = torch.zeros_like(d_cells)
    d_err = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    d_es = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    c.dev_build_neighbors(d_cells.data_ptr(), n, n, THR, R, LDIV)
    c.dev_step_rk(A, b, bs, 2e-5, d_cells.data_ptr(), n, n, d_new.data_ptr(), d_err.data_ptr(),
                  tol=1e-8, d_err_scaled=d_es.data_ptr())
    c.sync()
    torch.cuda.synchronize()
    new_ref, err_ref = orc.step_rk(A, b, bs, 2e-5, cells, nb, p)
    assert rel_err(d_new.cpu().numpy().T[:, 1:13], new_ref[:, 1:13]) <= 1e-10
    es_ref = orc.max_scaled_error(cells, err_ref, 1e-8)
    assert abs(d_es.item() - es_ref) <= 1e-10 * abs(es_ref)


def test_full_size_stage_velocities(orc):
    """The benchmark stage (binning -> list -> forces -> velocities) on the
    200,000-cell snapshot against the oracle's velocities on its own list."""
    import torch

    import bfm
    n = 200000
    cells, _ = bfm.make_dome(n, seed=42)
    p = bfm.core.make_params()
    c = bfm.Context(n, max_pairs=16 * n, params=p)
    d_cells = torch.from_numpy(np.ascontiguousarray(cells.T)).cuda()
    d_vel = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    c.dev_stage(d_cells.data_ptr(), n, n, THR, R, LDIV, d_vel.data_ptr())
    c.sync()
    torch.cuda.synchronize()
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    a2d = (cells[:, 6] < p.nz_threshold).astype(np.int32)
    v_ref = orc.get_velocities(cells, nb, a2d, None, p)
    err = rel_err(d_vel.cpu().numpy().T, v_ref)
    print(f"full-size stage rel_err={err:.3e}", flush=True)
    assert err <= 1e-10
