"""Host logic of the growth / division loop (no GPU): the division random
stream, divideCells (growth.hpp:456-720) in the product's host code against the
oracle's brute-force restatement, and closed-form pins of the division
geometry and the draw order."""
import numpy as np
import pytest

from gpu_common import random_box

R, RCELL, LDIV = 0.8, 0.5, 2.4


def _cfg(**kw):
    from bfm import sim
    return sim.sim_config(**kw)


def test_rng_matches_mt19937():
    """uniform_01 on mt19937: x * 2^-32 of the raw 32-bit outputs (numpy's
    MT19937 with the reference generator's seeding), and the C++ standard's
    known answer: the 10000th output for the default seed 5489 is 4123659995."""
    from bfm import sim
    for seed in (0, 1, 42, 255, 2**32 - 1):
        bg = np.random.MT19937()
        bg._legacy_seeding(seed)
        want = bg.random_raw(2000).astype(np.float64) / 2**32
        assert np.array_equal(sim.rng_uniform(seed, 2000), want)
    u = sim.rng_uniform(5489, 10000)
    assert u[-1] == 4123659995 / 2**32


def _population(n, side, seed, n_div, zmin=0.75, tilt=0.0):
    """Surface cells in a box, the first n_div of them past the division length."""
    cells = random_box(n, side, seed, nz_min=0.0, zmin=0.0)
    rng = np.random.default_rng(seed + 1000)
    cells[:, 3] = zmin + rng.uniform(0, 0.1, n)
    v = np.column_stack([rng.normal(size=(n, 2)), tilt * np.abs(rng.normal(size=n))])
    cells[:, 4:7] = v / np.linalg.norm(v, axis=1)[:, None]
    hl = rng.uniform(0.2, 1.15, n)
    hl[:n_div] = rng.uniform(1.21, 1.5, n_div)
    cells[:, 13], cells[:, 14] = 2 * hl, hl
    cells[:, 16] = rng.uniform(1.2, 1.8, n)
    cells[:, 20] = rng.integers(1, 3, n)
    return np.asfortranarray(cells)


@pytest.mark.parametrize("seed,side,tilt", [(1, 40.0, 0.0), (2, 12.0, 0.0), (3, 8.0, 0.3),
                                            (4, 6.0, 1.0), (5, 25.0, 0.5)])
def test_divide_matches_oracle(orc, seed, side, tilt):
    """Product host division == oracle division, bit for bit, over sparse and
    crowded populations (crowding exercises the no-room and retry paths)."""
    from bfm import sim
    cfg = _cfg(daughter_angle_xy_bound=0.3, daughter_angle_z_bound=0.2, growth_mean=(1.5, 1.0),
               growth_std=(0.15, 0.3))
    cells = _population(60, side, seed, n_div=12, tilt=tilt)
    got, par, tries = sim.divide_cells(cfg, cells, 0.37, seed)
    want, par_o, tries_o = orc.divide_cells(cfg, cells, 0.37, seed)
    assert got.shape == want.shape == (72, 22)
    assert tries == tries_o
    assert np.array_equal(got, want)
    assert np.array_equal(par, par_o)


def test_divide_retry_path_taken(orc):
    """A crowded case where some daughters fail the 2 Rcell criterion: more
    than one try, and product and oracle still agree."""
    from bfm import sim
    cfg = _cfg(daughter_angle_xy_bound=1.5, daughter_angle_z_bound=0.0)
    found = False
    for seed in range(10):
        cells = _population(40, 15.0, 100 + seed, n_div=20)
        got, _, tries = sim.divide_cells(cfg, cells, 0.0, seed)
        want, _, tries_o = orc.divide_cells(cfg, cells, 0.0, seed)
        assert tries == tries_o and np.array_equal(got, want)
        found |= tries > 1
    assert found


def test_division_geometry_closed_form(orc):
    """One isolated horizontal cell, no re-orientation: the daughters are
    collinear with the mother, their lengths add to l - 2R, the outer poles of
    the colony stay where the mother's were, the gap between the daughters'
    segment ends is 2R, ids and lineage follow growth.hpp:700-712."""
    cfg = _cfg(daughter_angle_xy_bound=0.0, daughter_angle_z_bound=0.0)
    cells = np.zeros((1, 22), order="F")
    th = 0.7
    n = np.array([np.cos(th), np.sin(th), 0.0])
    r = np.array([3.0, -2.0, 0.75])
    cells[0, 1:4], cells[0, 4:7] = r, n
    cells[0, 13], cells[0, 14] = 2.5, 1.25
    cells[0, 16], cells[0, 20], cells[0, 21] = 1.5, 1, 100
    out, par, tries = orc.divide_cells(cfg, cells, 2.0, 7)
    assert tries == 1 and out.shape[0] == 2
    l1, l2 = out[0, 13], out[1, 13]
    assert abs(l1 + l2 - (2.5 - 2 * R)) < 1e-14
    for row in out:
        assert np.allclose(row[4:7], n, atol=0, rtol=0)
    back_mother = r - (1.25 + R) * n
    front_mother = r + (1.25 + R) * n
    assert np.allclose(out[0, 1:4] - (l1 / 2 + R) * n, back_mother, atol=1e-13)
    assert np.allclose(out[1, 1:4] + (l2 / 2 + R) * n, front_mother, atol=1e-13)
    gap = (out[1, 1:4] - l2 / 2 * n) - (out[0, 1:4] + l1 / 2 * n)
    assert abs(np.linalg.norm(gap) - 2 * R) < 1e-13
    assert out[0, 0] == 1 and out[1, 0] == 2 and list(par) == [-1, 0, 0]
    assert out[0, 15] == 2.0 and out[1, 15] == 2.0


def test_division_draw_order(orc):
    """Draw order of divideCells for one cell with room to rotate: four angle
    draws (xy1, z1, xy2, z2), the length ratio (Box-Muller on the next two),
    then the mother's and the daughter's growth rates (two each). Pinned by
    closed forms: a horizontal mother rotated out of plane by theta has
    n_z = sin(theta); M = l1 / (l - 2R)."""
    from bfm import sim
    bxy, bz, sd = 0.0, 0.4, 0.125
    cfg = _cfg(daughter_angle_xy_bound=bxy, daughter_angle_z_bound=bz, daughter_length_std=sd,
               growth_mean=(1.5, 1.5), growth_std=(0.15, 0.15))
    cells = np.zeros((1, 22), order="F")
    cells[0, 1:4], cells[0, 4:7] = [0, 0, 0.75], [1, 0, 0]
    cells[0, 13], cells[0, 14] = 2.5, 1.25
    cells[0, 16], cells[0, 20], cells[0, 21] = 1.5, 1, 100
    seed = 2024
    out, _, _ = orc.divide_cells(cfg, cells, 0.0, seed)
    u = sim.rng_uniform(seed, 10)

    def normal(a, b):
        return np.sqrt(-2 * np.log(a)) * np.cos(2 * np.pi * b)

    z1 = -bz + 2 * bz * u[1]
    z2 = -bz + 2 * bz * u[3]
    assert abs(out[0, 6] - np.sin(z1)) < 1e-15 and abs(out[1, 6] - np.sin(z2)) < 1e-15
    M = min(max(0.5 + sd * normal(u[4], u[5]), 0.0), 1.0)
    assert abs(out[0, 13] / (2.5 - 2 * R) - M) < 1e-14
    assert abs(out[0, 16] - (1.5 + 0.15 * normal(u[6], u[7]))) < 1e-14
    assert abs(out[1, 16] - (1.5 + 0.15 * normal(u[8], u[9]))) < 1e-14
    # the product host code consumes the same stream
    got, _, _ = sim.divide_cells(cfg, cells, 0.0, seed)
    assert np.array_equal(got, out)


def test_no_room_keeps_orientation(orc):
    """A dividing cell with a neighbor closer than 2 Rcell is not re-oriented
    (growth.hpp:534-545): both daughters keep the mother's orientation even
    with non-zero angle bounds."""
    cfg = _cfg(daughter_angle_xy_bound=0.5, daughter_angle_z_bound=0.5)
    cells = np.zeros((2, 22), order="F")
    cells[:, 0] = [0, 1]
    cells[0, 1:4], cells[0, 4:7] = [0, 0, 0.75], [1, 0, 0]
    cells[1, 1:4], cells[1, 4:7] = [0, 0.9, 0.75], [1, 0, 0]
    cells[:, 13] = [2.5, 1.0]
    cells[:, 14] = cells[:, 13] / 2
    cells[:, 16], cells[:, 20], cells[:, 21] = 1.5, 1, 100
    out, par, tries = orc.divide_cells(cfg, cells, 0.0, 3)
    assert tries == 1 and out.shape[0] == 3
    assert np.array_equal(out[0, 4:7], [1, 0, 0]) and np.array_equal(out[2, 4:7], [1, 0, 0])
    assert list(par) == [-1, -1, 0, 0]
