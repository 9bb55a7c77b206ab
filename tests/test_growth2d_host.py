"""Host logic of the 2-D tree's growth loop (no GPU): divideCells of ABM-2d
(growth.hpp:438-697) in the product's host code against the oracle's
restatement, closed-form pins of the division geometry and of the draw order,
the isotropic-JKR adhesion of the 2-D composition tied to the pinned 3-D JKR
pair force, and invariants of the oracle's loop."""
import numpy as np
import pytest

R, RCELL, L0 = 0.8, 0.5, 0.4
LDIV = 2 * L0 + 2 * R


def _cfg(**kw):
    from bfm.model2d import sim_config_2d
    return sim_config_2d(**kw)


def _population(n, seed, n_div, spacing=(3.0, 1.4)):
    """A 2-D layer in rows with the first n_div cells past the division length."""
    from bfm.model2d import random_layer
    c = random_layer(n, seed, spacing=spacing)
    rng = np.random.default_rng(seed + 77)
    lvals = rng.uniform(LDIV + 0.01, LDIV + 0.5, n_div)
    c[:n_div, 9] = lvals
    c[:n_div, 10] = lvals / 2
    c[:, 16] = rng.integers(1, 3, n)
    return c


@pytest.mark.parametrize("seed,n,n_div,bound,spacing", [
    (0, 40, 5, 0.05, (3.0, 1.4)),
    (1, 60, 12, 0.5, (3.0, 1.4)),
    (2, 80, 20, 1.5, (2.6, 1.15)),
    (3, 30, 30, 0.05, (3.2, 1.6)),
    (4, 120, 7, 3.0, (2.5, 1.1)),
])
def test_divide_matches_oracle(orc, seed, n, n_div, bound, spacing):
    """The product's divideCells (bfm2d_divide_cells) equals the oracle's bit
    for bit: cells, lineage and number of placement tries."""
    from bfm.model2d import divide_cells_2d
    cfg = _cfg(daughter_angle_bound=bound)
    cells = _population(n, seed, n_div, spacing)
    got, par, tries = divide_cells_2d(cfg, cells, 0.37, seed + 11)
    want, par_o, tries_o = orc.divide_cells_2d(cfg, cells, 0.37, seed + 11)
    assert got.shape == want.shape == (n + n_div, cells.shape[1])
    assert np.array_equal(got, want)
    assert np.array_equal(par, par_o) and tries == tries_o


def test_divide_retries_are_exercised(orc):
    """A crowded layer with a large re-orientation bound needs more than one
    placement try, and the product follows the oracle through them."""
    from bfm.model2d import divide_cells_2d
    seen = []
    for bound in (0.6, 1.0, 3.0):
        cfg = _cfg(daughter_angle_bound=bound)
        for seed in range(4):
            cells = _population(150, seed, 6, spacing=(5.0, 1.4))
            got, _, tries = divide_cells_2d(cfg, cells, 0.0, seed)
            want, _, tries_o = orc.divide_cells_2d(cfg, cells, 0.0, seed)
            assert np.array_equal(got, want) and tries == tries_o
            seen.append(tries)
    assert min(seen) < 10 and max(seen) == 10 and len(set(seen)) > 3


def _lone_cell(l=2.9, angle=0.4, group=1, growth=1.5):
    from bfm.model2d import NCOLS2D
    c = np.zeros((1, NCOLS2D))
    c[0, 0] = 0
    c[0, 1:3] = (1.3, -0.7)
    c[0, 3:5] = (np.cos(angle), np.sin(angle))
    c[0, 9], c[0, 10] = l, l / 2
    c[0, 12], c[0, 13], c[0, 14], c[0, 15], c[0, 16] = growth, 0.7, 720.0, 720.0, group
    return c


def test_division_geometry_without_rotation(orc):
    """With no re-orientation the daughters lie on the mother's axis: lengths
    M (l - 2R) and (1 - M)(l - 2R), the rear tip of daughter 1 and the front
    tip of daughter 2 are the mother's tips, the inner caps touch at a
    center-line gap of 2R, birth times t, ids 1 and 2 with parent 0."""
    cfg = _cfg(daughter_angle_bound=0.0)
    c = _lone_cell()
    out, par, tries = orc.divide_cells_2d(cfg, c, 1.25, 9)
    assert out.shape[0] == 2 and tries == 1
    r, n, hl = c[0, 1:3], c[0, 3:5], c[0, 10]
    m, d = out[0], out[1]
    assert np.allclose(m[3:5], n, atol=0) and np.allclose(d[3:5], n, atol=0)
    assert abs(m[9] + d[9] - (c[0, 9] - 2 * R)) < 1e-14
    rear = r - (hl + R) * n
    front = r + (hl + R) * n
    assert np.allclose(m[1:3] - (m[10] + R) * n, rear, atol=1e-14)
    assert np.allclose(d[1:3] + (d[10] + R) * n, front, atol=1e-14)
    gap = np.dot(d[1:3] - m[1:3], n) - m[10] - d[10]
    assert abs(gap - 2 * R) < 1e-14
    assert m[11] == d[11] == 1.25
    assert (m[0], d[0]) == (1, 2) and list(par) == [-1, 0, 0]


def test_division_draw_order(orc):
    """A dividing cell with room takes, in order: two re-orientation angles
    (uniform on [-b, b]), the length ratio (0.5 + sigma N(0, 1), Box-Muller
    from two uniforms), the mother's growth rate, then the daughter's (each
    mean + std N(0, 1)) -- eight uniforms of the stream."""
    from bfm import sim
    from bfm.model2d import divide_cells_2d
    b, sd, seed = 0.2, 0.125, 31
    cfg = _cfg(daughter_angle_bound=b, daughter_length_std=sd, growth_mean=1.5, growth_std=0.15)
    c = _lone_cell()
    out, _, tries = divide_cells_2d(cfg, c, 0.0, seed)
    assert tries == 1
    u = sim.rng_uniform(seed, 8)

    def normal(a, bb):
        return np.sqrt(-2 * np.log(a)) * np.cos(2 * np.pi * bb)
    th1, th2 = -b + 2 * b * u[0], -b + 2 * b * u[1]
    M = min(max(0.5 + sd * normal(u[2], u[3]), 0.0), 1.0)
    a0 = np.arctan2(c[0, 4], c[0, 3])
    assert abs(np.arctan2(out[0, 4], out[0, 3]) - (a0 + th1)) < 1e-12
    assert abs(np.arctan2(out[1, 4], out[1, 3]) - (a0 + th2)) < 1e-12
    assert abs(out[0, 9] / (c[0, 9] - 2 * R) - M) < 1e-12
    assert abs(out[0, 12] - (1.5 + 0.15 * normal(u[4], u[5]))) < 1e-12
    assert abs(out[1, 12] - (1.5 + 0.15 * normal(u[6], u[7]))) < 1e-12


def test_no_room_means_no_rotation(orc):
    """A dividing cell closer than 2 Rcell to another cell is not re-oriented
    and draws no angles: the length ratio comes from the first two uniforms."""
    from bfm import sim
    from bfm.model2d import divide_cells_2d
    cfg = _cfg(daughter_angle_bound=1.0, daughter_length_std=0.125)
    c = np.vstack([_lone_cell(angle=0.0), _lone_cell(l=1.0, angle=0.0)])
    c[1, 0] = 1
    c[1, 1:3] = c[0, 1:3] + (0.0, 0.9)          # parallel, 0.9 < 2 Rcell apart
    out, _, tries = divide_cells_2d(cfg, c, 0.0, 5)
    want, _, _ = orc.divide_cells_2d(cfg, c, 0.0, 5)
    assert np.array_equal(out, want) and tries == 1
    assert np.array_equal(out[0, 3:5], c[0, 3:5]) and np.array_equal(out[2, 3:5], c[0, 3:5])
    u = sim.rng_uniform(5, 2)
    M = 0.5 + 0.125 * np.sqrt(-2 * np.log(u[0])) * np.cos(2 * np.pi * u[1])
    assert abs(out[0, 9] / (c[0, 9] - 2 * R) - M) < 1e-12


def test_divide_capacity_and_arguments():
    from bfm.model2d import _lib, divide_cells_2d
    import ctypes as C
    cfg = _cfg()
    c = _population(10, 0, 3)
    with pytest.raises(Exception):
        divide_cells_2d(cfg, c[:, :12], 0.0, 1)         # fewer than 17 columns
    lib = _lib()
    buf = np.zeros((11, c.shape[1]), order="F")
    buf[:10] = c
    n = C.c_int(10)
    par = np.full(40, -1, dtype=np.int32)
    npar = C.c_int(10)
    rc = lib.bfm2d_divide_cells(C.cast(C.byref(cfg), C.c_void_p), buf.ctypes.data_as(C.POINTER(C.c_double)),
                                C.byref(n), 11, c.shape[1], 0.0, 1, par.ctypes.data_as(C.POINTER(C.c_int)),
                                C.byref(npar), 40, None)
    assert rc == 2 and n.value == 10                      # BFM_ERR_CAPACITY, nothing changed


# ---- the isotropic-JKR adhesion of the 2-D composition ---------------------------
@pytest.mark.parametrize("delta", [0.01, 0.12, 0.3, 0.55])
@pytest.mark.parametrize("gamma", [10.0, 100.0, 1000.0])
def test_jkr_isotropic_2d_is_the_3d_jkr_force(orc, delta, gamma):
    """For a pair in the plane the 2-D adhesive dE/dq of mode 4 equals the
    3-D JKR pair force and constrained torques of the oracle's 3-D pair loop
    (pinned there against finite differences of the JKR energy) on the same
    pair at z = 0: in-plane position and orientation components."""
    from bfm.model2d import make_params, ADHESION2D_JKR_ISOTROPIC, NCOLS2D
    from bfm.core import Params, ADHESION_JKR_ISOTROPIC, JKR_EXACT
    rng = np.random.default_rng(int(delta * 1000 + gamma))
    th1, th2 = rng.uniform(0, np.pi, 2)
    n1, n2 = np.array([np.cos(th1), np.sin(th1)]), np.array([np.cos(th2), np.sin(th2)])
    c2 = np.zeros((2, NCOLS2D))
    c2[:, 0] = (0, 1)
    c2[0, 1:3], c2[0, 3:5] = (0.0, 0.0), n1
    c2[1, 3:5] = n2
    c2[:, 9], c2[:, 10] = (1.4, 1.0), (0.7, 0.5)
    c2[:, 13], c2[:, 14], c2[:, 16] = 0.7, 720.0, 1
    # place cell 1 so that the closest points are 2R - delta apart
    c2[1, 1:3] = (0.3, 2.5)
    for _ in range(4):          # row (1, 0): d points from cell 1 to cell 0
        nb = orc.get_cell_neighbors_2d(c2, 5.0, R, LDIV)
        d = nb[0, 2:4]
        c2[1, 1:3] += d / np.linalg.norm(d) * (np.linalg.norm(d) - (2 * R - delta))
    nb = orc.get_cell_neighbors_2d(c2, 5.0, R, LDIV)
    assert nb.shape[0] == 1 and abs(2 * R - np.linalg.norm(nb[0, 2:4]) - delta) < 1e-9
    p2 = make_params(adhesion_mode=ADHESION2D_JKR_ISOTROPIC, gamma=gamma, strength=1.0)
    _, adh = orc.forces_2d(c2, nb, p2, np.ones(1, dtype=np.int32))
    c3 = np.zeros((2, 22))
    c3[:, 0] = (0, 1)
    c3[:, 1:3], c3[:, 4:6] = c2[:, 1:3], c2[:, 3:5]
    c3[:, 13], c3[:, 14] = c2[:, 9], c2[:, 10]
    c3[:, 20], c3[:, 21] = 1, gamma
    nb3 = orc.get_cell_neighbors(c3, 5.0, R, LDIV)
    p3 = Params()
    p3.R, p3.Rcell, p3.E0 = R, RCELL, 3900.0
    p3.adhesion_mode, p3.jkr_mode, p3.colidx_gamma, p3.ncols = ADHESION_JKR_ISOTROPIC, JKR_EXACT, 21, 22
    p3.no_surface = 1
    F = orc.cell_cell_forces_ic(c3, nb3, p3)
    scale = np.max(np.abs(F))
    assert np.max(np.abs(adh[:, 0:2] - F[:, 0:2])) <= 1e-12 * scale
    assert np.max(np.abs(adh[:, 2:4] - F[:, 3:5])) <= 1e-12 * scale
    assert np.max(np.abs(F[:, [2, 5]])) <= 1e-12 * scale


def test_jkr_isotropic_2d_needs_to_adhere(orc):
    """Rows without to_adhere carry no adhesion in mode 4 (the 2-D loop's gate)."""
    from bfm.model2d import make_params, random_layer, ADHESION2D_JKR_ISOTROPIC
    c = random_layer(60, 3)
    nb = orc.get_cell_neighbors_2d(c, 2 * R + LDIV, R, LDIV)
    p2 = make_params(adhesion_mode=ADHESION2D_JKR_ISOTROPIC)
    _, adh = orc.forces_2d(c, nb, p2, np.zeros(nb.shape[0], dtype=np.int32))
    assert not np.any(adh)
    _, adh1 = orc.forces_2d(c, nb, p2, np.ones(nb.shape[0], dtype=np.int32))
    assert np.any(adh1)


# ---- the oracle's loop -------------------------------------------------------------
def test_oracle_growth_loop_invariants(orc):
    """Founder to 16 cells: ids 0..2(n-1) with the lineage of every division,
    unit orientations, t = sum of the accepted dt, a run that repeats exactly."""
    from bfm.model2d import founder_cell_2d, make_params, ADHESION2D_JKR_ISOTROPIC
    cfg = _cfg(n_cells=16, seed=42)
    p = make_params(adhesion_mode=ADHESION2D_JKR_ISOTROPIC)
    amap = np.array([[0, 0], [0, 1]], dtype=np.int32)
    cells, par, tr, it, steps = orc.simulate_2d(cfg, p, amap, founder_cell_2d(), cap=64)
    n = cells.shape[0]
    assert n >= 16 and it == len(tr) and steps >= 2 * it
    assert len(par) == 1 + 2 * (n - 1)
    ids = np.sort(cells[:, 0].astype(int))
    live = set(range(len(par))) - set(par[par >= 0])
    assert set(ids) == live
    assert np.allclose(np.hypot(cells[:, 3], cells[:, 4]), 1.0, atol=1e-14)
    assert np.all(cells[:, 9] <= LDIV + 0.05) and np.all(cells[:, 9] > 0)
    cells2, par2, tr2, it2, _ = orc.simulate_2d(cfg, p, amap, founder_cell_2d(), cap=64)
    assert np.array_equal(cells, cells2) and np.array_equal(tr, tr2) and np.array_equal(par, par2)


def test_oracle_first_division_time(orc):
    """A lone founder grows by growCells' Euler step of dl/dt = g (4R/3 + l),
    so l + 4R/3 gains a factor 1 + g dt per step at dt = max_stepsize (no
    forces, zero error): the loop stops right after the first division, at
    the first step count k with (L0 + 4R/3)(1 + g dt)^k > Ldiv + 4R/3."""
    from bfm.model2d import founder_cell_2d, make_params
    cfg = _cfg(n_cells=2, seed=3)
    p = make_params()
    cells, par, tr, it, _ = orc.simulate_2d(cfg, p, None, founder_cell_2d(), cap=8)
    dt, g = cfg.max_stepsize, 1.5
    assert np.all(tr == dt)
    k = int(np.ceil(np.log((LDIV + 4 * R / 3) / (L0 + 4 * R / 3)) / np.log1p(g * dt)))
    # grown k times, divided at the top of iteration k, stepped once more
    assert it in (k + 1, k + 2)
    assert cells.shape[0] == 2 and list(par) == [-1, 0, 0]
