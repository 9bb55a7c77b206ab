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
