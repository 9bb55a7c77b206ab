"""Oracle geometry pinned by the reference's known answers and by brute force.

distBetweenCells known answers: ABM-3d/ABM-3d-cpp/test/src/testDistBetweenCells.cpp:143-506
(tolerance 1e-8 as there). The skew branch (CGAL in the reference) is also pinned by
brute-force minimisation over (s, t) and by symmetry invariants.
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _kats():
    with open(os.path.join(GOLD, "dist_kats.json")) as f:
        return json.load(f)


def test_nearest_cell_body_coord_kats(orc):
    for c in _kats()["nearest"]:
        s = orc.nearest_cell_body_coord(c["r"], c["n"], c["half_l"], c["q"])
        assert abs(s - c["target"]) < 1e-8


@pytest.mark.parametrize("case", _kats()["cases"], ids=lambda c: c["name"])
def test_dist_between_cells_kats(orc, case):
    d, s, t = orc.dist_between_cells(case["r1"], case["n1"], case["half_l1"], case["r2"],
                                     case["n2"], case["half_l2"])
    assert np.allclose(d, case["d"], atol=1e-8, rtol=0), (d, case["d"])
    assert abs(s - case["s"]) < 1e-8 and abs(t - case["t"]) < 1e-8, (s, t)


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def _brute(r1, n1, h1, r2, n2, h2):
    """Grid search then local refinement of |r2 + t n2 - r1 - s n1| over the box."""
    ss = np.linspace(-h1, h1, 401)
    ts = np.linspace(-h2, h2, 401)
    S, T = np.meshgrid(ss, ts, indexing="ij")
    D = (r2[None, None] + T[..., None] * n2 - r1[None, None] - S[..., None] * n1)
    k = np.argmin((D ** 2).sum(-1))
    s0, t0 = S.flat[k], T.flat[k]
    ds, dt = ss[1] - ss[0], ts[1] - ts[0]
    for _ in range(60):
        ss = np.linspace(max(-h1, s0 - ds), min(h1, s0 + ds), 41)
        ts = np.linspace(max(-h2, t0 - dt), min(h2, t0 + dt), 41)
        S, T = np.meshgrid(ss, ts, indexing="ij")
        D = (r2[None, None] + T[..., None] * n2 - r1[None, None] - S[..., None] * n1)
        k = np.argmin((D ** 2).sum(-1))
        s0, t0 = S.flat[k], T.flat[k]
        ds, dt = ds / 10, dt / 10
        if ds < 1e-13 and dt < 1e-13:
            break
    return np.sqrt(((r2 + t0 * n2 - r1 - s0 * n1) ** 2).sum()), s0, t0


def test_skew_branch_matches_brute_force(orc):
    rng = np.random.default_rng(3)
    for _ in range(60):
        r1, r2 = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        n1, n2 = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        if abs(n1 @ n2) >= 0.9999:
            continue
        h1, h2 = rng.uniform(0.2, 1.6, 2)
        d, s, t = orc.dist_between_cells(r1, n1, h1, r2, n2, h2)
        dist_b, sb, tb = _brute(r1, n1, h1, r2, n2, h2)
        assert abs(np.linalg.norm(d) - dist_b) < 1e-9
        # the minimiser is unique for non-parallel segments
        assert abs(s - sb) < 1e-6 and abs(t - tb) < 1e-6
        assert np.allclose(d, r2 + t * n2 - r1 - s * n1, atol=1e-12)


def test_distance_symmetries(orc):
    rng = np.random.default_rng(5)
    for _ in range(200):
        r1, r2 = rng.uniform(-2, 2, 3), rng.uniform(-2, 2, 3)
        n1, n2 = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        h1, h2 = rng.uniform(0.2, 1.6, 2)
        if abs(n1 @ n2) >= 0.9999:
            continue   # the near-parallel averaging heuristic is not symmetric
        d, s, t = orc.dist_between_cells(r1, n1, h1, r2, n2, h2)
        d2, s2, t2 = orc.dist_between_cells(r2, n2, h2, r1, n1, h1)
        d3, s3, t3 = orc.dist_between_cells(r1, -n1, h1, r2, n2, h2)
        # unique minimiser: swapping the cells and reversing n1 map exactly
        assert abs(np.linalg.norm(d) - np.linalg.norm(d2)) < 1e-12
        assert abs(np.linalg.norm(d) - np.linalg.norm(d3)) < 1e-12
        assert np.allclose(d2, -d, atol=1e-12) and abs(s2 - t) < 1e-9 and abs(t2 - s) < 1e-9
        assert abs(s3 + s) < 1e-9 and abs(t3 - t) < 1e-9


def test_near_parallel_branch_distance_upper_bounds_minimum(orc):
    """The near-parallel branch averages endpoint vectors (distances.hpp:227-337);
    it always returns a vector between two points of the segments."""
    rng = np.random.default_rng(11)
    for _ in range(200):
        n1 = _unit(rng.normal(size=3))
        n2 = _unit(n1 + rng.normal(scale=3e-3, size=3)) * rng.choice([-1, 1])
        r1, r2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        h1, h2 = rng.uniform(0.2, 1.6, 2)
        d, s, t = orc.dist_between_cells(r1, n1, h1, r2, n2, h2)
        assert -h1 - 1e-12 <= s <= h1 + 1e-12 and -h2 - 1e-12 <= t <= h2 + 1e-12
        assert np.allclose(d, r2 + t * n2 - r1 - s * n1, atol=1e-12)
        dist_b, _, _ = _brute(r1, n1, h1, r2, n2, h2)
        assert np.linalg.norm(d) >= dist_b - 1e-9


def test_neighbor_list_brute_force_equals_binned(orc):
    import bfm
    for n, seed in ((1500, 1), (4000, 2)):
        cells, _ = bfm.make_dome(n, seed=seed)
        a = orc.get_cell_neighbors(cells, 2.0, 0.8, 2.4)
        b = orc.get_cell_neighbors(cells, 2.0, 0.8, 2.4, binned=True)
        assert np.array_equal(a, b)
        i, j = a[:, 0], a[:, 1]
        assert np.all(j < i)
        key = i * n + j
        assert np.all(np.diff(key) > 0)   # rows sorted (i asc, j asc), no duplicates
        assert np.all(np.linalg.norm(a[:, 2:5], axis=1) < 2.0)


def test_neighbor_list_small_edge_cases(orc):
    import bfm
    cells, _ = bfm.make_dome(10, seed=3)
    assert orc.get_cell_neighbors(cells[:0], 2.0, 0.8, 2.4).shape == (0, 7)
    assert orc.get_cell_neighbors(cells[:1], 2.0, 0.8, 2.4).shape == (0, 7)
