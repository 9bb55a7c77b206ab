"""The 2-D model in the oracle (ABM-2d/ABM-2d-cpp: getCellNeighbors,
cellCellRepulsiveForces, cellCellAdhesiveForces, getVelocities,
stepRungeKuttaAdaptive; mechanics.hpp:293-374, 448-694, 1503-1817 of ABM-2d),
pinned by:

* the reference's own distance cases (test/src/testDistBetweenCells.cpp:
  113-344 of ABM-2d, to 1e-8 as there) through the oracle's closest-point
  computation on cells at z = 0 (the reference's non-parallel branch calls
  CGAL's segment distance, which is not available; the closest points of two
  non-parallel segments are unique, so any exact method agrees to rounding);
* brute force on random pairs: the distance against a dense search over both
  centerlines;
* the forces as gradients of the reference's potentials (potentialHertz,
  potentialJKR, potentialKihara, adhesion.hpp:70-177 of ABM-2d): central
  differences in the positions, and in the orientations after the projection
  onto the unit-norm constraint, as test/src/testHertz.cpp:195-270 checks;
* the velocities as -dE/dq over the composite drag l (eta0 + eta1 scd / R)
  and l^3 (...) / 12;
* the RK step: with torque-free contacts the stages coincide and every
  tableau gives cells + dt v with zero error; with torques the stage
  orientations enter (a stage computed here from getVelocities agrees).
"""
import numpy as np
import pytest

R, RCELL, E0, ECELL = 0.8, 0.5, 3900.0, 3.9e6

# (r1x, r1y, n1x, n1y, hl1, r2x, r2y, n2x, n2y, hl2, dx, dy, s, t) of
# testDistBetweenCells.cpp (the targets evaluated as written there)
_S = -0.07692307692
KATS = [
    (2, 1, 0.6, 0.8, 1, 0, 3, 0, 1, 0.5, -2, 1.5, 0, -0.5),
    (4, 4, 5 / 13, 12 / 13, 2.7, 5, 3, 0, 1, 0.5,
     (5 + 0.5 * 0) - (4 + _S * 5 / 13), (3 + 0.5 * 1) - (4 + _S * 12 / 13), _S, 0.5),
    (4, 4, 5 / 13, 12 / 13, 2.7, 5, 3, 0, 1, 5.0, 0.0, 0.0, 2.6, 3.4),
    (4, 4, 5 / 13, 12 / 13, 1.0, 5, 3, 0, 1, 5.0,
     (5 + 1.9230769231 * 0) - (4 + 1.0 * 5 / 13), (3 + 1.9230769231) - (4 + 12 / 13), 1.0, 1.9230769231),
    (4, 4, 5 / 13, 12 / 13, 0.6, 6, 2, 0, 1, 0.5,
     (6 + 0.5 * 0) - (4 - 0.6 * 5 / 13), (2 + 0.5) - (4 - 0.6 * 12 / 13), -0.6, 0.5),
    (0, 0, 1, 0, 1.0, -2, 2, 1, 0, 0.75, (-2 + 0.75) - (0 - 1), 2, -1, 0.75),
    (0, 0, 1, 0, 1.0, -0.8, 2, 1, 0, 1.5, 0, 2, -0.15, 0.65),
    (0, 0, 1, 0, 1.0, -0.8, 2, -1, 0, 1.5, 0, 2, -0.15, -0.65),
    (0, 0, 1, 0, 1.0, 0.1, 2, 1, 0, 1.5, 0, 2, 0, -0.1),
    (0, 0, 1, 0, 1.0, 0.1, 2, 1, 0, 0.6, 0, 2, 0.1, 0),
    (0, 0, 1, 0, 1.0, 0.5, 2, 1, 0, 1.1, 0, 2, 0.2, -0.3),
    (0, 0, 1, 0, 1.0, 0.5, 2, -1, 0, 1.1, 0, 2, 0.2, 0.3),
    (0.5, 2, 1, 0, 1.1, 0, 0, 1, 0, 1.0, 0, -2, -0.3, 0.2),
    (0, 0, 1, 0, 1.0, 5, 2, 1, 0, 1.5, (5 - 1.5) - 1, 2, 1, -1.5),
    (0, 0, -1, 0, 1.0, 5, 2, 1, 0, 1.5, (5 - 1.5) - 1, 2, -1, -1.5),
    (0, 0, 1, 0, 1.0, 5, 0, 1, 0, 1.5, (5 - 1.5) - 1, 0, 1, -1.5),
]


@pytest.mark.parametrize("case", range(len(KATS)))
def test_distance_matches_reference_cases(orc, case):
    r1x, r1y, n1x, n1y, hl1, r2x, r2y, n2x, n2y, hl2, dx, dy, s, t = KATS[case]
    d, s_, t_ = orc.dist_between_cells([r1x, r1y, 0], [n1x, n1y, 0], hl1, [r2x, r2y, 0], [n2x, n2y, 0], hl2)
    assert abs(d[0] - dx) < 1e-8 and abs(d[1] - dy) < 1e-8 and d[2] == 0.0
    assert abs(s_ - s) < 1e-8 and abs(t_ - t) < 1e-8


def _layer(n, seed):
    from bfm.model2d import random_layer
    return random_layer(n, seed)


def test_neighbor_rows_against_brute_force(orc):
    """For non-parallel pairs (the branch where the reference calls CGAL)
    every row's distance is the minimum over both centerlines (dense grid, to
    its resolution) and its contact points reproduce its distance vector; no
    pair below the threshold is missed. (Nearly parallel pairs, |n_i . n_j| >=
    0.9999, take the reference's averaged contact points by design: the
    reference cases above cover that branch.)"""
    c = _layer(60, 1)
    thr, Ldiv = 2.0, 2.4
    nb = orc.get_cell_neighbors_2d(c, thr, R, Ldiv)
    assert nb.shape[1] == 6 and nb.shape[0] > 40
    u = np.linspace(-1, 1, 401)
    found = set()
    for row in nb:
        i, j = int(row[0]), int(row[1])
        found.add((i, j))
        assert j < i
        if abs(np.dot(c[i, 3:5], c[j, 3:5])) >= 0.9999:
            continue
        p = c[i, 1:3] + np.outer(u * c[i, 10], c[i, 3:5])
        q = c[j, 1:3] + np.outer(u * c[j, 10], c[j, 3:5])
        dmin = np.min(np.linalg.norm(p[:, None, :] - q[None, :, :], axis=2))
        dist = np.hypot(row[2], row[3])
        assert dist <= dmin + 1e-12
        assert dist >= dmin - 2 * (2.0 / 400) * max(c[i, 10], c[j, 10])   # grid resolution
        # the row's contact points reproduce its distance vector
        ci = c[i, 1:3] + row[4] * c[i, 3:5]
        cj = c[j, 1:3] + row[5] * c[j, 3:5]
        np.testing.assert_allclose(cj - ci, row[2:4], atol=1e-12)
    for i in range(60):
        for j in range(i):
            if (i, j) in found:
                continue
            d, _, _ = orc.dist_between_cells(np.r_[c[i, 1:3], 0], np.r_[c[i, 3:5], 0], c[i, 10],
                                             np.r_[c[j, 1:3], 0], np.r_[c[j, 3:5], 0], c[j, 10])
            assert np.linalg.norm(d) >= thr


def _pot_hertz(dist):
    if dist <= 2 * RCELL:
        d0 = (4 * R + 6 * RCELL) / 5.0
        return 2.5 * E0 * np.sqrt(R) * (2 * (R - RCELL)) ** 1.5 * (d0 - dist) + \
            ECELL * np.sqrt(RCELL) * (2 * RCELL - dist) ** 2.5
    if dist <= 2 * R:
        return E0 * np.sqrt(R) * (2 * R - dist) ** 2.5
    return 0.0


def _pot_jkr(dist, dmin):
    if dist <= dmin:
        d0 = (2 * R + dmin) / 2.0
        return -np.pi * R * (2 * R - dmin) * (d0 - dist)
    if dist <= 2 * R:
        return -0.5 * np.pi * R * (2 * R - dist) ** 2
    return 0.0


def _pot_kihara(dist, e, dmin):
    if dist <= dmin:
        t1 = e * dist * (1 / (dmin ** e * dmin) - 1 / ((2 * R) ** e * 2 * R))
        return t1 + (e + 1) * (1 / (2 * R) ** e - 1 / dmin ** e)
    if dist <= 2 * R:
        return -1 / dist ** e - e * dist / ((2 * R) ** e * 2 * R) + (e + 1) / (2 * R) ** e
    return 0.0


def _energy(orc, q, hl, pot):
    r1, n1, r2, n2 = q[0:2], q[2:4], q[4:6], q[6:8]
    d, _, _ = orc.dist_between_cells(np.r_[r1, 0], np.r_[n1, 0], hl[0], np.r_[r2, 0], np.r_[n2, 0], hl[1])
    return pot(np.linalg.norm(d))


@pytest.mark.parametrize("kind,dist", [("hertz", 1.3), ("hertz", 0.9), ("jkr", 1.3), ("jkr", 1.05),
                                       ("kihara", 1.3), ("kihara", 1.1)])
def test_forces_are_potential_gradients(orc, kind, dist):
    from bfm.model2d import make_params, ADHESION2D_JKR, ADHESION2D_KIHARA, ADHESION2D_NONE
    mode = {"hertz": ADHESION2D_NONE, "jkr": ADHESION2D_JKR, "kihara": ADHESION2D_KIHARA}[kind]
    dmin, e, strength = 1.2, 6.0, 1.7
    p = make_params(adhesion_mode=mode, strength=strength, mindist=dmin, distance_exp=e, cap_exponent=1.5,
                    E0=E0, Ecell=ECELL)
    rng = np.random.default_rng(7)
    hl = np.array([1.0, 0.7])
    th = rng.uniform(0, 2 * np.pi, 2)
    n1, n2 = np.array([np.cos(th[0]), np.sin(th[0])]), np.array([np.cos(th[1]), np.sin(th[1])])
    # place cell 2 so that its nearest point is at distance `dist` from cell 1's point s = 0.3
    perp = np.array([-n1[1], n1[0]])
    r1 = np.array([1.0, 2.0])
    r2 = r1 + 0.3 * n1 + dist * perp - 0.2 * n2
    c = np.zeros((2, 17), order="F")
    c[:, 0] = [0, 1]
    c[0, 1:3], c[0, 3:5], c[1, 1:3], c[1, 3:5] = r1, n1, r2, n2
    c[:, 9], c[:, 10] = 2 * hl, hl
    c[:, 13], c[:, 14] = 0.7, 720.0
    d, s, t = orc.dist_between_cells(np.r_[r1, 0], np.r_[n1, 0], hl[0], np.r_[r2, 0], np.r_[n2, 0], hl[1])
    nb = np.asfortranarray([[1, 0, -d[0], -d[1], t, s]])       # the reference's row order: i > j
    rep, adh = orc.forces_2d(c, nb, p, to_adhere=[1])
    dE = rep + adh
    if kind == "hertz":
        pot = _pot_hertz
        assert not adh.any()
    elif kind == "jkr":
        pot = lambda x: _pot_hertz(x) + strength * _pot_jkr(x, dmin)           # noqa: E731
    else:
        pot = lambda x: _pot_hertz(x) + strength * _pot_kihara(x, e, dmin)     # noqa: E731
    q0 = np.r_[r1, n1, r2, n2]
    h = 1e-6
    g = np.zeros(8)
    for k in range(8):
        qp, qm = q0.copy(), q0.copy()
        qp[k] += h
        qm[k] -= h
        g[k] = (_energy(orc, qp, hl, pot) - _energy(orc, qm, hl, pot)) / (2 * h)
    for cidx, off, nn in ((0, 0, n1), (1, 4, n2)):
        gr, gn = g[off:off + 2], g[off + 2:off + 4]
        gn_c = gn - np.dot(nn, gn) * nn          # the unit-norm constraint
        scale = max(np.abs(g).max(), 1.0)
        np.testing.assert_allclose(dE[cidx, 0:2], gr, rtol=0, atol=2e-6 * scale)
        np.testing.assert_allclose(dE[cidx, 2:4], gn_c, rtol=0, atol=2e-6 * scale)
    assert np.abs(dE).max() > 1.0
    np.testing.assert_allclose(dE[0, :2], -dE[1, :2], atol=1e-9 * np.abs(dE).max())


def test_velocities_are_forces_over_drag(orc):
    from bfm.model2d import make_params, ADHESION2D_JKR
    p = make_params(adhesion_mode=ADHESION2D_JKR, strength=0.3)
    c = _layer(80, 3)
    nb = orc.get_cell_neighbors_2d(c, 2.0, R, 2.4)
    ta = (np.arange(nb.shape[0]) % 3 != 0).astype(np.int32)
    noise = np.asfortranarray(np.random.default_rng(4).uniform(-5, 5, (80, 4)))
    rep, adh = orc.forces_2d(c, nb, p, ta)
    v = orc.get_velocities_2d(c, nb, p, ta, noise)
    assert np.isfinite(v).all()
    drag = c[:, 13] + c[:, 14] * p.surface_contact_density / R
    K, L = c[:, 9] * drag, c[:, 9] ** 3 * drag / 12
    dE = rep + adh + noise
    np.testing.assert_allclose(v[:, :2], -dE[:, :2] / K[:, None], rtol=1e-13)
    np.testing.assert_allclose(v[:, 2:], -dE[:, 2:] / L[:, None], rtol=1e-13)
    # to_adhere switches the adhesion of a row off
    _, adh_all = orc.forces_2d(c, nb, p, np.ones(nb.shape[0], np.int32))
    assert np.abs(adh_all - adh).max() > 1e-3


def test_rk_with_torque_free_contacts(orc):
    """Pairs touching at both centers (s = t = 0) exert no torque, so every
    stage sees the same forces: cells + dt v for any tableau, zero error."""
    import bfm
    from bfm.model2d import make_params
    p = make_params()
    c = np.zeros((2, 17), order="F")
    c[:, 0] = [0, 1]
    c[0, 1:5] = [0, 0, 1, 0]
    c[1, 1:5] = [0, 1.3, 0.6, 0.8]
    c[:, 9], c[:, 10] = 2.0, 1.0
    c[:, 13], c[:, 14] = 0.7, 720.0
    nb = np.asfortranarray([[1, 0, 0.0, -1.3, 0.0, 0.0]])
    v = orc.get_velocities_2d(c, nb, p)
    assert np.abs(v[:, :2]).max() > 0 and not v[:, 2:].any()
    for name in ("heun_euler", "bogacki_shampine", "dormand_prince"):
        A, b, bs, _ = bfm.tableau(name)
        new, err = orc.step_rk_2d(A, b, bs, 1e-3, c, nb, p)
        np.testing.assert_allclose(new[:, 1:3], c[:, 1:3] + 1e-3 * v[:, :2], rtol=0, atol=1e-15)
        assert np.abs(err).max() < 1e-15
        np.testing.assert_allclose(new[:, 5:7], v[:, :2], rtol=1e-14)


def test_rk_stages_use_stage_orientations(orc):
    """Two stages of Heun-Euler written out with getVelocities: the second at
    cells + dt k0 with the orientation renormalised (no flip in 2-D)."""
    import bfm
    from bfm.model2d import make_params
    p = make_params()
    c = _layer(40, 5)
    nb = orc.get_cell_neighbors_2d(c, 2.0, R, 2.4)
    dt = 2e-4
    k0 = orc.get_velocities_2d(c, nb, p)
    c1 = c.copy(order="F")
    c1[:, 1:5] += k0 * dt
    c1[:, 3:5] /= np.linalg.norm(c1[:, 3:5], axis=1)[:, None]
    k1 = orc.get_velocities_2d(c1, nb, p)
    A, b, bs, _ = bfm.tableau("heun_euler")
    new, err = orc.step_rk_2d(A, b, bs, dt, c, nb, p)
    want = c[:, 1:5] + (0.5 * k0 + 0.5 * k1) * dt
    want[:, 2:4] /= np.linalg.norm(want[:, 2:4], axis=1)[:, None]
    np.testing.assert_allclose(new[:, 1:5], want, rtol=0, atol=1e-13)
    np.testing.assert_allclose(err, (0.5 * k0 + 0.5 * k1) * dt - k0 * dt, rtol=0, atol=1e-15)
    assert np.abs(k1[:, 2:] - k0[:, 2:]).max() > 1e-6 * np.abs(k0[:, 2:]).max()   # torques moved
