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


@pytest.mark.parametrize("mode", ["hertz", "jkr_exact", "gamma_zero_mix", "table_1d", "table_2d", "surrogate"])
def test_stage_pair_forces_fused_into_compaction(orc, dome, mode):
    """bfm_dev_stage writes the isotropic pair-force records inside the
    neighbor-list compaction (Hertz, exact JKR, tables); the learned radius
    keeps the separate kernels. The stage's velocities equal, bit for bit, a
    separate build followed by getVelocities on the device list (which runs
    the pair-force kernel), and the oracle's velocities within 1e-10 (the
    learned radius: within its 1e-4 force tolerance of the exact oracle)."""
    import torch

    import bfm
    from bfm import surrogate as sur
    from bfm.jkr import reference_table, reference_table_2d
    cells, nb = dome
    cells = cells.copy(order="F")
    table = None
    if mode == "hertz":
        p = bfm.core.make_params(adhesion=False, ncols=21)
        cells = np.asfortranarray(cells[:, :21])
    elif mode == "jkr_exact":
        p = bfm.core.make_params()
    elif mode == "gamma_zero_mix":
        cells[::3, 21] = 0.0
        p = bfm.core.make_params()
    elif mode == "table_1d":
        p = bfm.core.make_params(jkr_mode=bfm.core.JKR_TABLE_1D)
        table = reference_table(R, RCELL, 3900.0, 100.0)
    elif mode == "table_2d":
        rng = np.random.default_rng(22)
        cells[:, 21] = np.where(rng.random(cells.shape[0]) < 0.5, 50.0, 500.0)
        ov, gm, rad = reference_table_2d(R, RCELL, 3900.0, 500.0)
        p = bfm.core.make_params(jkr_mode=bfm.core.JKR_TABLE_2D)
        table = (ov, np.ascontiguousarray(rad.ravel()), gm)
    else:
        p = bfm.core.make_params(jkr_mode=bfm.core.JKR_SURROGATE)
    n = cells.shape[0]
    print(f"fused stage mode={mode} n={n}", flush=True)
    c = bfm.Context(n, max_pairs=16 * n, params=p)
    if table is not None:
        c.set_jkr_table(*table)
        orc.set_jkr_table(*table)
    if mode == "surrogate":
        sur.install(c, sur.load())
    d_cells = torch.from_numpy(np.ascontiguousarray(cells.T)).cuda()
    d_v1 = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    d_v2 = torch.zeros((6, n), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    c.dev_stage(d_cells.data_ptr(), n, n, THR, R, LDIV, d_v1.data_ptr())
    c.dev_build_neighbors(d_cells.data_ptr(), n, n, THR, R, LDIV)
    c.dev_get_velocities(d_cells.data_ptr(), n, n, d_v2.data_ptr())
    c.sync()
    torch.cuda.synchronize()
    v1, v2 = d_v1.cpu().numpy().T, d_v2.cpu().numpy().T
    assert np.array_equal(v1, v2)
    p_ref = bfm.core.make_params(adhesion=False, ncols=21) if mode == "hertz" else (
        bfm.core.make_params() if mode == "surrogate" else p)
    v_ref = orc.get_velocities(cells, nb, _a2d(cells, p), None, p_ref)
    err = rel_err(v1, v_ref)
    print(f"rel_err={err:.3e}", flush=True)
    assert err <= (1e-4 if mode == "surrogate" else TOL)
