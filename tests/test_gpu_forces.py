"""GPU conservative forces and velocities (getConservativeForces / getVelocities,
mechanics.hpp:1702-1769, 2011-2347) against the oracle on identical inputs.
Bar: normwise per-column relative error <= 1e-10 (max_i |v - v_ref| / max_i |v_ref|).
Cell-cell sums run in the reference's row order on both sides, so the only
differences come from powers (x*sqrt(x) on the GPU vs libm pow in the oracle)."""
import numpy as np
import pytest

from gpu_common import LDIV, R, RCELL, THR, random_box, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


@pytest.fixture(scope="module")
def dome():
    import bfm
    from oracle import oracle as orc
    cells, _ = bfm.make_dome(12000, seed=21)
    nb = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    return cells, nb


def _a2d(cells, p):
    return (cells[:, 6] < p.nz_threshold).astype(np.int32)


def _check(orc, cells, nb, p, noise=None, table=None):
    import bfm
    c = bfm.Context(max(cells.shape[0], 16), max_pairs=max(nb.shape[0], 16), params=p)
    if table is not None:
        c.set_jkr_table(*table)
        orc.set_jkr_table(*table)
    a2d = _a2d(cells, p)
    v = c.get_velocities(cells, nb, a2d, noise)
    v_ref = orc.get_velocities(cells, nb, a2d, noise, p)
    err = rel_err(v, v_ref)
    print(f"rel_err={err:.3e}", flush=True)
    assert err <= TOL
    f = c.get_conservative_forces(cells, nb, a2d)
    f_ref = orc.conservative_forces(cells, nb, a2d, p)
    assert rel_err(f, f_ref) <= TOL
    return v


def test_jkr_exact(orc, dome):
    import bfm
    cells, nb = dome
    _check(orc, cells, nb, bfm.core.make_params())


def test_hertz_only(orc, dome):
    import bfm
    cells, nb = dome
    _check(orc, cells, nb, bfm.core.make_params(adhesion=False, ncols=21))


def test_gamma_zero_cells_use_hertz_with_separate_hard_core(orc, dome):
    import bfm
    cells, nb = dome
    cells = cells.copy(order="F")
    cells[::3, 21] = 0.0
    _check(orc, cells, nb, bfm.core.make_params())


def test_table_1d(orc, dome):
    import bfm
    from bfm.jkr import reference_table
    cells, nb = dome
    p = bfm.core.make_params(jkr_mode=bfm.core.JKR_TABLE_1D)
    _check(orc, cells, nb, p, table=reference_table(R, RCELL, 3900.0, 100.0))


def test_table_2d_two_phenotypes(orc, dome):
    import bfm
    from bfm.jkr import reference_table_2d
    cells, nb = dome
    cells = cells.copy(order="F")
    rng = np.random.default_rng(22)
    cells[:, 21] = np.where(rng.random(cells.shape[0]) < 0.5, 50.0, 500.0)
    ov, gm, rad = reference_table_2d(R, RCELL, 3900.0, 500.0)
    p = bfm.core.make_params(jkr_mode=bfm.core.JKR_TABLE_2D)
    _check(orc, cells, nb, p, table=(ov, np.ascontiguousarray(rad.ravel()), gm))


def test_no_surface_and_noise(orc, dome):
    import bfm
    cells, nb = dome
    rng = np.random.default_rng(23)
    noise = np.asfortranarray(rng.uniform(-50, 50, (cells.shape[0], 6)))
    _check(orc, cells, nb, bfm.core.make_params(no_surface=True), noise=noise)


def test_coulomb_off_and_horizontal_cells(orc):
    """Horizontal cells (assume_2d) on the surface and tilted contacting cells."""
    import bfm
    from oracle import oracle as orc_
    cells = random_box(6000, 40.0, 24, zmin=0.3)
    cells[::4, 6] = 0.0
    xy = cells[::4, 4:6]
    cells[::4, 4:6] = xy / np.linalg.norm(xy, axis=1)[:, None]
    nb = orc_.get_cell_neighbors(cells, THR, R, LDIV)
    for coeff in (0.0, 1.0):
        p = bfm.core.make_params(cell_surface_coulomb_coeff=coeff)
        assert _a2d(cells, p).sum() >= 1000
        _check(orc, cells, nb, p)


def test_unsorted_neighbor_rows(orc, dome):
    """Rows in arbitrary order: both sides accumulate in the given row order."""
    import bfm
    cells, nb = dome
    perm = np.random.default_rng(25).permutation(nb.shape[0])
    _check(orc, cells, np.asfortranarray(nb[perm]), bfm.core.make_params())


def test_empty_neighbor_list(orc, dome):
    import bfm
    cells, _ = dome
    _check(orc, cells, np.zeros((0, 7), order="F"), bfm.core.make_params())
