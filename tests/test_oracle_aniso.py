"""The anisotropic JKR table path in the oracle (forcesAnisotropicJKRLagrange,
table overloads, ABM-3d/ABM-3d-cpp/include/adhesion.hpp:734-1046; used through
jkr_data->getJKRForces, mechanics.hpp:1468-1479), pinned independently of its
own formulas:

* nearestValue (utils.hpp:1169-1215) against a brute-force argmin whose ties
  go to the upper grid point (the reference's `<` test), duplicates included;
* the lookup against a restatement built on other formulas: angles from
  arctan2(|a x b|, a . b) folded with min(x, pi - x) instead of acos and the
  reference's branch, grid indices from a brute-force argmin, on tables whose
  entries encode their own index (so a wrong, transposed or swapped index
  shows); samples within 1e-9 of a Voronoi boundary are skipped;
* the assembly into the cell-cell forces: one pair gives dE/dr_i = F d/|d|,
  dE/dr_j = -F d/|d| and torques s_i F d/|d|, -s_j F d/|d| (include_constraint
  false), plus the separately added hard-core term past 2 (R - Rcell).
"""
import numpy as np
import pytest

R, RCELL = 0.8, 0.5


def _nearest_brute(v, x):
    """The nearer of the last grid point <= x and the first one > x, the
    upper one on a tie; the end points outside the grid."""
    v = np.asarray(v)
    if x <= v[0]:
        return 0
    if x >= v[-1]:
        return len(v) - 1
    lo = int(np.nonzero(v <= x)[0].max())
    hi = int(np.nonzero(v > x)[0].min())
    return hi if v[hi] - x <= x - v[lo] else lo


def test_nearest_value_ties_go_up(orc):
    v = np.array([0.0, 1.0, 1.0, 2.0, 4.0, 4.0, 4.0, 7.0])
    xs = np.concatenate([np.arange(-1.0, 8.5, 0.5), [1.0, 4.0, 5.5, 3.0]])
    for x in xs:
        k = orc.nearest_value(v, x)
        # a tie between the neighbors below and above goes up; in a run of
        # equal grid values the one taken is the last at or below x or the
        # first above it
        assert abs(v[k] - x) == np.min(np.abs(v - x)), x
        assert k == _nearest_brute(v, x), x
    assert [orc.nearest_value(v, x) for x in (0.5, 1.0, 1.5, 3.0, 4.0, 5.5)] == [1, 2, 3, 4, 6, 7]


def _angle(a, b):
    return np.arctan2(np.linalg.norm(np.cross(a, b)), np.dot(a, b))


def _fold(x):
    return min(x, np.pi - x)


def _coded_tables(seed, gammas=None):
    """Tables whose force entry is its own flat index and whose Rx entries are
    random points of the Rx range."""
    from bfm.core import AnisoJKRTables
    rng = np.random.default_rng(seed)
    theta = np.sort(rng.uniform(0, np.pi / 2, 9))
    half_l = np.sort(rng.uniform(0.3, 1.8, 6))
    coords = np.sort(rng.uniform(0, 1, 7))
    rx = np.sort(rng.uniform(0.8, 2.4, 8))
    phi = np.sort(rng.uniform(0, np.pi / 2, 5))
    delta = np.sort(rng.uniform(0, 0.6, 11))
    curv = rng.uniform(0.8, 2.4, (9, 6, 7))
    ng = 1 if gammas is None else len(gammas)
    n = 8 * 8 * 5 * 11 * ng
    fr = np.stack([np.arange(n, dtype=np.float64), -np.arange(n, dtype=np.float64)], axis=-1)
    shape = (8, 8, 5, 11) + (() if gammas is None else (ng,)) + (2,)
    return AnisoJKRTables(theta, half_l, coords, curv, rx, phi, delta, fr.reshape(shape),
                          gamma=None if gammas is None else np.asarray(gammas, float))


def _near_boundary(v, x, eps=1e-9):
    v = np.asarray(v)
    mids = 0.5 * (v[1:] + v[:-1])
    return np.any(np.abs(mids - x) < eps)


@pytest.mark.parametrize("variable_gamma", [False, True])
def test_lookup_against_independent_restatement(orc, variable_gamma):
    gammas = [20.0, 60.0, 90.0, 150.0] if variable_gamma else None
    t = _coded_tables(3, gammas)
    orc.set_aniso_tables(t)
    rng = np.random.default_rng(4)
    checked = 0
    for _ in range(3000):
        n1 = rng.normal(size=3); n1 /= np.linalg.norm(n1)
        n2 = rng.normal(size=3); n2 /= np.linalg.norm(n2)
        d = rng.normal(size=3); d *= rng.uniform(1.0, 1.7) / np.linalg.norm(d)
        hl1, hl2 = rng.uniform(0.2, 2.0, 2)
        s, tt = rng.uniform(-hl1, hl1), rng.uniform(-hl2, hl2)
        gamma = rng.uniform(10, 170)
        dist = np.linalg.norm(d)
        if dist >= 2 * R:           # no contact
            f, a = orc.aniso_jkr_force(n1, hl1, n2, hl2, d, R, gamma, s, tt, 2 * (R - RCELL))
            assert (f, a) == (0.0, 0.0)
            continue
        delta = min(2 * R - dist, 2 * (R - RCELL))
        dn = d / dist
        th1 = _fold(_angle(dn, n1))
        th2 = _fold(_angle(-dn, n2))
        ph = _fold(_angle(n1, n2))
        q1 = (th1, t.theta), (hl1, t.half_l), (abs(s) / hl1, t.coords)
        q2 = (th2, t.theta), (hl2, t.half_l), (abs(tt) / hl2, t.coords)
        if any(_near_boundary(g, x) for x, g in q1 + q2):
            continue
        i1 = [_nearest_brute(g, x) for x, g in q1]
        i2 = [_nearest_brute(g, x) for x, g in q2]
        rx1, rx2 = sorted([t.curvature_rx[tuple(i1)], t.curvature_rx[tuple(i2)]])
        rest = [(rx1, t.rx), (rx2, t.rx), (ph, t.phi), (delta, t.delta)]
        if variable_gamma:
            rest.append((gamma, t.gamma))
        if any(_near_boundary(g, x) for x, g in rest):
            continue
        idx = tuple(_nearest_brute(g, x) for x, g in rest)
        want = t.force_radius[idx]
        f, a = orc.aniso_jkr_force(n1, hl1, n2, hl2, d, R, gamma, s, tt, 2 * (R - RCELL))
        assert (f, a) == (want[0], want[1]), (idx, f)
        checked += 1
    assert checked > 1500


@pytest.mark.parametrize("ic", [False, True])
def test_pair_assembly(orc, ic):
    from bfm.core import make_params
    from bfm.jkr import synthetic_aniso_tables
    t = synthetic_aniso_tables()
    orc.set_aniso_tables(t)
    p = make_params(anisotropic=True)
    rng = np.random.default_rng(9)
    for dist in (1.3, 0.95):            # inside the EPS, and past 2 (R - Rcell) = 0.6
        cells = np.zeros((2, 22), order="F")
        cells[:, 0] = [0, 1]
        n = rng.normal(size=(2, 3)); n[:, 2] = np.abs(n[:, 2])
        cells[:, 4:7] = n / np.linalg.norm(n, axis=1)[:, None]
        cells[:, 13], cells[:, 14] = [2.0, 1.6], [1.0, 0.8]
        cells[:, 21] = 100.0
        cells[0, 1:4] = [0, 0, 2]
        d = rng.normal(size=3); d *= dist / np.linalg.norm(d)
        cells[1, 1:4] = cells[0, 1:4] + d
        nb = np.asfortranarray([[0, 1, d[0], d[1], d[2], 0.3, -0.2]], dtype=np.float64)
        F, _ = orc.aniso_jkr_force(cells[0, 4:7], 1.0, cells[1, 4:7], 0.8, d, R, 100.0, 0.3, -0.2,
                                   2 * (R - RCELL))
        assert F != 0.0
        over = 2 * R - dist
        hard = p.repulsion_prefactors[1] * max(over - 2 * (R - RCELL), 0.0) ** 1.5
        dn = d / dist
        dE = (orc.cell_cell_forces_ic if ic else orc.cell_cell_forces)(cells, nb, p)
        np.testing.assert_allclose(dE[0, :3], (F + hard) * dn, rtol=1e-14, atol=1e-12)
        np.testing.assert_allclose(dE[1, :3], -(F + hard) * dn, rtol=1e-14, atol=1e-12)
        if not ic:
            np.testing.assert_allclose(dE[0, 3:], 0.3 * (F + hard) * dn, rtol=1e-14, atol=1e-12)
            np.testing.assert_allclose(dE[1, 3:], 0.2 * (F + hard) * dn, rtol=1e-14, atol=1e-12)
        else:
            # constrained torques are perpendicular to the orientation
            for c in range(2):
                assert abs(np.dot(dE[c, 3:], cells[c, 4:7])) < 1e-12 * np.abs(dE[c, 3:]).max()
