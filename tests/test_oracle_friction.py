"""Cell-cell kinetic friction (cellCellFrictionForces, ABM-3d/ABM-3d-cpp/include/
mechanics.hpp:1031-1175) in the oracle, pinned by hand-computed contacts and by
physics the reference's formula must satisfy:

* two parallel cells side by side, one sliding: F = -eta sqrt(Req delta) v_t,
  the normal part of the relative velocity removed;
* Coulomb's cap: |F| = mu (4/3) E0 sqrt(Req) delta^1.5 with the direction kept;
* the 2-D switch near the surface (|d_z| / |d| < 1e-3 and both contact points
  below 1.01 R): the z parts of d and of the relative velocity are dropped;
* Newton's third law (the forces of a pair cancel);
* dissipation: the power of the friction forces against the stored
  velocities, sum_i -dE/dq_i . (dr_i, dn_i), equals -sum eta sqrt(Req delta)
  |v_t|^2 <= 0 (a wrong sign or s-factor of a torque breaks the identity);
* friction adds to the conservative cell-cell forces as a separate array.
"""
import numpy as np
import pytest

R = 0.8


def _params(friction=True, coulomb=1.0, no_surface=False):
    from bfm.core import make_params
    return make_params(friction=friction, cell_cell_coulomb_coeff=coulomb, no_surface=no_surface)


def _pair(rz, dy=1.5, dz=0.0, x_shift=0.0, v_i=(1.0, 0.3, 0.0), eta=(2.0, 3.0)):
    cells = np.zeros((2, 23), order="F")
    cells[:, 0] = [0, 1]
    cells[0, 1:4] = [0.0, 0.0, rz]
    cells[1, 1:4] = [x_shift, dy, rz + dz]
    cells[:, 4] = 1.0                       # both along x
    cells[:, 13], cells[:, 14] = 2.0, 1.0
    cells[:, 16], cells[:, 17], cells[:, 18], cells[:, 19], cells[:, 20] = 1.5, 0.7, 720, 100, 1
    cells[:, 21] = 100.0
    cells[:, 22] = eta
    cells[0, 7:10] = v_i
    return cells


def test_sliding_pair_by_hand(orc):
    cells = _pair(rz=5.0)
    nb = orc.get_cell_neighbors(cells, 2.0, R, 2.4)
    assert nb.shape[0] == 1
    f = orc.friction_forces(cells, nb, _params())
    # overlap 0.1, Req 0.4, eta = min(2, 3): F_1 = -2 sqrt(0.04) (1, 0, 0) = (-0.4, 0, 0)
    row_i, row_j = int(nb[0, 0]), int(nb[0, 1])
    np.testing.assert_allclose(f[row_i, :3], [0.4, 0, 0] if row_i == 0 else [-0.4, 0, 0], atol=1e-14)
    np.testing.assert_allclose(f[row_j, :3], -f[row_i, :3], atol=1e-14)
    # equal x extents: the contact coordinate is 0, no torque
    np.testing.assert_allclose(f[:, 3:], 0, atol=1e-14)


def test_offset_pair_torque(orc):
    """j shifted by 0.5 along x: the near-parallel branch takes the middle of
    the overlapping extents, s_0 = 0.25 and s_1 = -0.25; torque = s F."""
    cells = _pair(rz=5.0, x_shift=0.5)
    nb = orc.get_cell_neighbors(cells, 2.0, R, 2.4)
    f = orc.friction_forces(cells, nb, _params())
    s = {int(nb[0, 0]): nb[0, 5], int(nb[0, 1]): nb[0, 6]}
    assert s[0] == pytest.approx(0.25) and s[1] == pytest.approx(-0.25)
    # the force on cell 0 is -0.4 e_x (dE/dq = +0.4), on cell 1 +0.4 e_x
    np.testing.assert_allclose(f[0, :3], [0.4, 0, 0], atol=1e-14)
    np.testing.assert_allclose(f[0, 3:], [0.25 * 0.4, 0, 0], atol=1e-14)
    np.testing.assert_allclose(f[1, 3:], [-0.25 * -0.4, 0, 0], atol=1e-14)


def test_coulomb_cap(orc):
    cells = _pair(rz=5.0, v_i=(1e4, 0.0, 2e3))
    nb = orc.get_cell_neighbors(cells, 2.0, R, 2.4)
    f = orc.friction_forces(cells, nb, _params(coulomb=0.5))
    cap = 0.5 * (4 / 3) * 3900 * np.sqrt(0.4) * 0.1 ** 1.5
    fi = f[0, :3]
    assert np.linalg.norm(fi) == pytest.approx(cap, rel=1e-14)
    np.testing.assert_allclose(fi / np.linalg.norm(fi), [1e4, 0, 2e3] / np.linalg.norm([1e4, 0, 2e3]),
                               rtol=1e-14)
    # coefficient <= 0: no cap
    f = orc.friction_forces(cells, nb, _params(coulomb=0.0))
    assert np.linalg.norm(f[0, :3]) == pytest.approx(0.4 * np.linalg.norm([1e4, 0, 2e3]), rel=1e-14)


def test_two_d_switch_near_surface(orc):
    """d_z / |d| = 6.7e-4 < 1e-3: near the surface (z < 1.01 R) the z parts are
    dropped (overlap from the in-plane distance 1.5); high above it they are not."""
    low = _pair(rz=0.7, dz=0.001, v_i=(1.0, 0.0, 0.7))
    nb = orc.get_cell_neighbors(low, 2.0, R, 2.4)
    f = orc.friction_forces(low, nb, _params())
    np.testing.assert_allclose(f[0, :3], [0.4, 0, 0], atol=1e-14)
    high = _pair(rz=5.0, dz=0.001, v_i=(1.0, 0.0, 0.7))
    nb = orc.get_cell_neighbors(high, 2.0, R, 2.4)
    f = orc.friction_forces(high, nb, _params())
    assert abs(f[0, 2]) > 0.2           # the z velocity is resisted in 3-D
    # no surface: never 2-D
    f = orc.friction_forces(low, orc.get_cell_neighbors(low, 2.0, R, 2.4), _params(no_surface=True))
    assert abs(f[0, 2]) > 0.2


def _random_layer(n, seed):
    from gpu_common import random_box
    cells = random_box(n, 9.0, seed)
    c = np.zeros((n, 23), order="F")
    c[:, :22] = cells
    c[:, 22] = np.random.default_rng(seed + 1).uniform(0.5, 3.0, n)
    return c


@pytest.mark.parametrize("coulomb", [0.0, 1.0, 1e-3])
def test_third_law_and_dissipation(orc, coulomb):
    cells = _random_layer(300, 11)
    nb = orc.get_cell_neighbors(cells, 2.0, R, 2.4)
    # no surface: every pair in 3-D, so the identity below is exact
    f = orc.friction_forces(cells, nb, _params(coulomb=coulomb, no_surface=True))
    scale = np.max(np.abs(f[:, :3]))
    assert scale > 0
    assert np.max(np.abs(f[:, :3].sum(axis=0))) < 1e-11 * scale * len(cells)
    power = -np.sum(f[:, :3] * cells[:, 7:10]) - np.sum(f[:, 3:] * cells[:, 10:13])
    # -sum eta sqrt(Req delta) |v_t|^2 from the pairs, when nothing is capped
    want = 0.0
    for k in range(nb.shape[0]):
        i, j = int(nb[k, 0]), int(nb[k, 1])
        d = nb[k, 2:5]
        delta = 2 * R - np.linalg.norm(d)
        if delta <= 0:
            continue
        dn = d / np.linalg.norm(d)
        rv = (cells[i, 7:10] + nb[k, 5] * cells[i, 10:13]) - (cells[j, 7:10] + nb[k, 6] * cells[j, 10:13])
        vt = rv - rv.dot(dn) * dn
        want -= min(cells[i, 22], cells[j, 22]) * np.sqrt(0.4 * delta) * vt.dot(vt)
    assert power <= 0
    if coulomb == 0.0:
        assert power == pytest.approx(want, rel=1e-10)
    else:
        assert power >= want * (1 + 1e-10)      # the cap only removes dissipation


def test_friction_adds_to_conservative_forces(orc):
    cells = _random_layer(200, 5)
    nb = orc.get_cell_neighbors(cells, 2.0, R, 2.4)
    on = orc.cell_cell_forces(cells, nb, _params())
    off = orc.cell_cell_forces(cells, nb, _params(friction=False, ))
    fr = orc.friction_forces(cells, nb, _params())
    np.testing.assert_array_equal(on, off + fr)
