"""GPU neighbor search (getCellNeighbors, mechanics.hpp:520-593) against the oracle:
the lists must be identical, element for element (indices, distance vectors,
contact coordinates), including row order."""
import numpy as np
import pytest

from gpu_common import LDIV, R, THR, random_box

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import bfm
    return bfm.Context(220000, max_pairs=16 * 220000, params=bfm.core.make_params())


@pytest.mark.parametrize("n,seed", [(500, 1), (5000, 2), (20000, 3)])
def test_dome_lists_identical(ctx, orc, n, seed):
    import bfm
    cells, _ = bfm.make_dome(n, seed=seed)
    print(f"dome n={n} seed={seed}", flush=True)
    got = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV)
    assert got.shape == ref.shape and np.array_equal(got, ref)


def test_full_size_dome_identical_to_oracle(ctx, orc):
    """200,000 cells (the benchmark snapshot): identical to the oracle's list
    (the oracle's binned search, itself equal to its brute force on smaller inputs)."""
    import bfm
    cells, _ = bfm.make_dome(200000, seed=42)
    got = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    assert got.shape == ref.shape and np.array_equal(got, ref)
    key = got[:, 0] * 200000 + got[:, 1]
    assert np.all(np.diff(key) > 0) and np.all(got[:, 1] < got[:, 0])


def test_near_parallel_population(ctx, orc):
    """Nematic cells: most pairs take the near-parallel branch (|cos| >= 0.9999)."""
    cells = random_box(4000, 30.0, 7)
    rng = np.random.default_rng(8)
    v = np.array([0.6, 0.0, 0.8]) + rng.normal(scale=2e-3, size=(4000, 3))
    cells[:, 4:7] = v / np.linalg.norm(v, axis=1)[:, None]
    got = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV)
    assert np.array_equal(got, ref)
    n1 = cells[got[:, 0].astype(int), 4:7]
    n2 = cells[got[:, 1].astype(int), 4:7]
    assert np.mean(np.abs((n1 * n2).sum(1)) >= 0.9999) > 0.5


def test_long_cells_prefilter(ctx, orc):
    """Half-lengths up to 3.0: the reference prefilter |r_i - r_j| < thr + Ldiv + 2R
    drops pairs whose centerlines are within thr; the GPU must drop the same ones."""
    cells = random_box(3000, 25.0, 9)
    rng = np.random.default_rng(10)
    hl = rng.uniform(0.4, 3.0, 3000)
    cells[:, 13], cells[:, 14] = 2 * hl, hl
    got = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [0, 1, 2, 3])
def test_tiny_inputs(ctx, orc, n):
    cells = random_box(max(n, 1), 1.5, 11)[:n]
    got = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("bin_scale,block,variant", [
    (0.5, 256, 0), (2.0, 64, 0), (1.0, 512, 0), (0.34, 128, 0),
    (0.5, 0, 1), (2.0, 0, 1), (1.0, 0, 1), (0.34, 0, 1),
    (1.0, 0, 2), (0.5, 0, 3), (1.25, 0, 3),
    (1.0, 0, 4), (0.5, 0, 4), (0.34, 0, 4), (2.0, 0, 4), (1.25, 0, 5), (0.75, 0, 5),
    (1.0, 0, 6), (0.5, 0, 6), (1.25, 0, 6),
    (1.0, 0, 7), (0.5, 0, 7), (1.25, 0, 7), (0.34, 0, 7), (2.0, 0, 7),
    (1.0, 0, 8), (0.5, 0, 8), (0.75, 0, 8), (1.25, 0, 8), (0.34, 0, 8), (2.0, 0, 8)])
def test_tuning_knobs_do_not_change_the_list(orc, bin_scale, block, variant):
    import bfm
    cells, _ = bfm.make_dome(8000, seed=12)
    c = bfm.Context(8000, params=bfm.core.make_params())
    c.set_tuning(bin_scale, block, variant)
    got = c.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("variant", [3, 4, 5, 6, 7, 8])
def test_dense_packing_queue_overflow_path(orc, variant):
    """A packing dense enough that a group's survivors overflow the LDS queue
    in one pass, so the candidate kernels redo the chunk in home batches."""
    import bfm
    from gpu_common import random_box
    cells = random_box(6000, 9.0, 21)
    c = bfm.Context(6000, max_pairs=500 * 6000, params=bfm.core.make_params())
    c.set_tuning(0, 0, variant)
    got = c.get_cell_neighbors(cells, THR, R, LDIV)
    ref = orc.get_cell_neighbors(cells, THR, R, LDIV, binned=True)
    assert got.shape[0] > 50 * 6000
    assert np.array_equal(got, ref)


def test_update_neighbor_distances(ctx, orc):
    import bfm
    cells, _ = bfm.make_dome(6000, seed=13)
    nb = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    rng = np.random.default_rng(14)
    moved = cells.copy(order="F")
    moved[:, 1:4] += rng.normal(scale=0.05, size=(6000, 3))
    v = moved[:, 4:7] + rng.normal(scale=0.05, size=(6000, 3))
    moved[:, 4:7] = v / np.linalg.norm(v, axis=1)[:, None]
    got = nb.copy(order="F")
    ctx.update_neighbor_distances(moved, got)
    ref = nb.copy(order="F")
    orc.update_neighbor_distances(moved, ref)
    assert np.array_equal(got, ref)


def test_device_build_export_matches_host_api(ctx):
    import torch

    import bfm
    cells, _ = bfm.make_dome(10000, seed=15)
    host = ctx.get_cell_neighbors(cells, THR, R, LDIV)
    d_cells = torch.from_numpy(np.ascontiguousarray(cells.T)).cuda()
    torch.cuda.synchronize()
    m = ctx.dev_build_neighbors(d_cells.data_ptr(), 10000, 10000, THR, R, LDIV)
    out = torch.zeros((7, m), dtype=torch.float64, device="cuda")
    ctx.dev_export_neighbors(out.data_ptr(), m)
    ctx.sync()
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().T, host)
