"""The 2-D model on the GPU (bfm2d_get_cell_neighbors, bfm2d_get_velocities,
bfm2d_step_rk; getCellNeighbors, getVelocities and stepRungeKuttaAdaptive of
ABM-2d/ABM-2d-cpp/include/mechanics.hpp:293-374, 1503-1817) against the
oracle on identical inputs. Bars: the neighbor rows bit for bit (same rows,
same distances and coordinates); velocities and RK steps within 1e-10
normwise per column, errors within 1e-10 of the step's displacement; the
columns the step does not update copied exactly."""
import numpy as np
import pytest

from gpu_common import rel_err

pytestmark = pytest.mark.gpu
R, LDIV, THR = 0.8, 2.4, 2.0
TOL = 1e-10


def _layer(n, seed):
    from bfm.model2d import random_layer
    return random_layer(n, seed)


@pytest.fixture(scope="module")
def dense():
    from oracle import oracle as orc
    c = _layer(3000, 71)
    return c, orc.get_cell_neighbors_2d(c, THR, R, LDIV)


@pytest.mark.parametrize("n,seed", [(3000, 71), (800, 72), (2, 73), (1, 74)])
def test_neighbors_match_oracle(orc, n, seed):
    from bfm.model2d import Model2D, make_params
    c = _layer(n, seed)
    # parallel pairs (the 0.9999 branch) among the cells
    if n > 10:
        c[1::7, 3:5] = c[0:-1:7, 3:5][: len(c[1::7])]
    m2 = Model2D(max(n, 16), make_params())
    got = m2.getCellNeighbors(c, THR, R, LDIV)
    want = orc.get_cell_neighbors_2d(c, THR, R, LDIV)
    print(f"n={n} rows={want.shape[0]}", flush=True)
    assert got.shape == want.shape and np.array_equal(got, want)


MODES = {"none": dict(), "jkr": dict(adhesion_mode=1, strength=40.0, mindist=1.2),
         "kihara": dict(adhesion_mode=2, strength=50.0, mindist=1.1, distance_exp=6.0)}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("with_noise", [False, True])
def test_velocities_match_oracle(orc, dense, mode, with_noise):
    from bfm.model2d import Model2D, make_params
    c, nb = dense
    p = make_params(**MODES[mode])
    ta = (np.arange(nb.shape[0]) % 4 != 1).astype(np.int32)
    noise = np.asfortranarray(np.random.default_rng(75).uniform(-3, 3, (c.shape[0], 4))) if with_noise else None
    m2 = Model2D(c.shape[0], p, max_pairs=max(nb.shape[0], 16))
    v = m2.getVelocities(c, nb, ta, noise)
    v_ref = orc.get_velocities_2d(c, nb, p, ta, noise)
    assert np.isfinite(v_ref).all()
    print(f"{mode} noise={with_noise} rel_err={rel_err(v, v_ref):.3e}", flush=True)
    assert rel_err(v, v_ref) <= TOL
    assert np.abs(v_ref).max() > 0
    if mode != "none":       # the adhesion term is present in the compared velocities
        v_none = orc.get_velocities_2d(c, nb, make_params(), ta, noise)
        assert rel_err(v_none, v_ref) > 1e-6


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", ["heun_euler", "bogacki_shampine", "dormand_prince"])
def test_step_rk_matches_oracle(orc, dense, mode, name):
    import bfm
    from bfm.model2d import Model2D, make_params
    c, nb = dense
    p = make_params(**MODES[mode])
    ta = (np.arange(nb.shape[0]) % 5 != 2).astype(np.int32)
    A, b, bs, _ = bfm.tableau(name)
    dt = 1e-5
    m2 = Model2D(c.shape[0], p, max_pairs=max(nb.shape[0], 16))
    new, err = m2.stepRungeKuttaAdaptive(A, b, bs, c, nb, dt, ta)
    new_ref, err_ref = orc.step_rk_2d(A, b, bs, dt, c, nb, p, ta)
    e = rel_err(new[:, 1:9], new_ref[:, 1:9])
    print(f"{mode} {name} rel_err={e:.3e}", flush=True)
    assert e <= TOL
    keep = [0] + list(range(9, c.shape[1]))
    assert np.array_equal(new[:, keep], new_ref[:, keep])
    disp = np.max(np.abs(new_ref[:, 1:5] - c[:, 1:5]), axis=0)
    assert np.all(disp > 0)
    assert np.all(np.max(np.abs(err - err_ref), axis=0) <= 1e-10 * disp)


def test_device_list_drives_the_step(orc):
    """The GPU-built list fed back to the step gives the oracle's step on the
    oracle's list (the lists are identical)."""
    import bfm
    from bfm.model2d import Model2D, make_params
    c = _layer(1500, 77)
    p = make_params(adhesion_mode=1, strength=0.4, mindist=1.2)
    m2 = Model2D(1500, p)
    nb = m2.getCellNeighbors(c, THR, R, LDIV)
    A, b, bs, _ = bfm.tableau("bogacki_shampine")
    new, _ = m2.stepRungeKuttaAdaptive(A, b, bs, c, nb, 1e-5)
    new_ref, _ = orc.step_rk_2d(A, b, bs, 1e-5, c, orc.get_cell_neighbors_2d(c, THR, R, LDIV), p)
    assert rel_err(new[:, 1:9], new_ref[:, 1:9]) <= TOL


def _embed(c):
    """2-D cells in the 3-D layout at z = 0 (for the oracle's distance update)."""
    c3 = np.zeros((c.shape[0], 22), order="F")
    c3[:, 0] = c[:, 0]
    c3[:, 1:3], c3[:, 4:6], c3[:, 13], c3[:, 14] = c[:, 1:3], c[:, 3:5], c[:, 9], c[:, 10]
    return c3


def test_device_path_trajectory(orc):
    """bfm2d_dev_build_neighbors (to_adhere by the driver's group rule),
    bfm2d_dev_step_rk and bfm2d_dev_update_neighbor_distances on device
    arrays: the exported list equals the oracle's, and 8 Bogacki-Shampine
    steps with the distances updated between them stay within 1e-9 of the
    oracle's loop using the same rule on the host."""
    import torch

    import bfm
    from bfm.model2d import (Model2D, make_params, dev_build_neighbors, dev_step_rk,
                             dev_update_neighbor_distances, to_adhere_rule)
    n = 2000
    c = _layer(n, 78)
    c[:, 16] = np.random.default_rng(79).integers(1, 3, n)
    amap = np.array([[0, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=np.int32)
    p = make_params(adhesion_mode=1, strength=40.0, mindist=1.2)
    m2 = Model2D(n, p, max_pairs=8 * n)
    a = torch.from_numpy(np.ascontiguousarray(c.T)).cuda()
    bbuf = torch.empty_like(a)
    m = dev_build_neighbors(m2, a.data_ptr(), n, n, THR, R, LDIV, amap)
    nb_ref = orc.get_cell_neighbors_2d(c, THR, R, LDIV)
    exp = torch.zeros((7, max(m, 1)), dtype=torch.float64, device="cuda")
    m2.ctx._check(m2.ctx.lib.bfm_dev_export_neighbors(m2.ctx.h, exp.data_ptr(), max(m, 1)))
    torch.cuda.synchronize()
    nb_dev = exp.cpu().numpy().T[:m][:, [0, 1, 2, 3, 5, 6]]
    assert m == nb_ref.shape[0] and np.array_equal(nb_dev, nb_ref)
    ta = to_adhere_rule(c, nb_ref, R, 0.5, amap)
    assert 0 < ta.sum() < len(ta)
    A, b, bs, _ = bfm.tableau("bogacki_shampine")
    dt = 1e-5
    x_ref = c.copy(order="F")
    nb = nb_ref.copy(order="F")
    src, dst = a, bbuf
    for k in range(8):
        dev_step_rk(m2, A, b, bs, dt, src.data_ptr(), n, n, dst.data_ptr(), n)
        dev_update_neighbor_distances(m2, dst.data_ptr(), n, n)
        src, dst = dst, src
        x_ref, _ = orc.step_rk_2d(A, b, bs, dt, x_ref, nb, p, ta)
        nb7 = np.asfortranarray(np.c_[nb[:, :4], np.zeros(len(nb)), nb[:, 4:]])
        orc.update_neighbor_distances(_embed(x_ref), nb7)
        nb = np.asfortranarray(nb7[:, [0, 1, 2, 3, 5, 6]])
    torch.cuda.synchronize()
    got = src.cpu().numpy().T
    e = rel_err(got[:, 1:9], x_ref[:, 1:9])
    print(f"8-step device trajectory rel_err={e:.3e}", flush=True)
    assert e <= 1e-9
    assert np.abs(got[:, 1:3] - c[:, 1:3]).max() > 1e-6
