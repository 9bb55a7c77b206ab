"""The oracle's 3-D cell-cell forces (cellCellInteractionForces, both overloads,
mechanics.hpp:1209-1350 and 1390-1659) against finite differences of the pair
energy, taken in mpmath.

The pair energy is a function of the overlap delta = 2R - |d| only, with
d = d0 + (dr_j + t dn_j) - (dr_i + s dn_i): the distance vector of the
neighbor row moved with the four cell coordinates while the contact
coordinates s, t stay at the row's closest-point values. Its gradient with
respect to (r_i, n_i, r_j, n_j) is the reference's dE/dq without the
orientation constraint (force +-F d_hat, "torques" +s_i F d_hat, -s_j F d_hat).

Energies, written from the contact mechanics rather than from the oracle:

* Hertz sphere-sphere contact of radius R/2 and modulus E0 inside the EPS
  shell: U = (8/15) E0 sqrt(R/2) delta^(5/2) (force (4/3) E0 sqrt(R/2)
  delta^(3/2)); beyond the shell thickness 2 (R - Rcell) the shell force is
  held at its value there, i.e. the energy continues linearly;
* the hard core, of radius Rcell and modulus Ecell, once the overlap exceeds
  the shell: (8/15) Ecell sqrt(Rcell/2) (delta - 2 (R - Rcell))^(5/2);
* JKR with work of adhesion 2 gamma, equivalent radius Req = R/2: the contact
  radius a is the larger root of delta = a^2/Req - 2 sqrt(pi gamma a / E0),
  and U(a) = 8 E0 a^5 / (15 Req^2) - (8/3) sqrt(pi gamma E0) a^(7/2) / Req
  + 2 pi gamma a^2 (dU/d delta = 4 E0 a^3 / (3 Req) - 4 sqrt(pi gamma E0 a^3),
  the JKR load); the overlap entering a is capped at the shell thickness
  (adhesion.hpp:195-196), so the energy continues linearly there too.

The repulsion prefactors handed to the oracle are the literal closed forms
computed here, not the package's helper, and that helper is checked against
the same closed forms separately. Planted mistakes (a doubled prefactor, a
flipped torque sign, swapped contact coordinates) must fail the comparison.
"""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40

R, RCELL, E0, ECELL = 0.8, 0.5, 3900.0, 3.9e6
REQ = R / 2
DC = 2 * (R - RCELL)                        # shell thickness, the EPS cap
H = mp.mpf("1e-12")                         # central-difference step (mpmath)
TOL = 1e-8                                  # relative, per component block


def _closed_form_prefactors():
    """(4/3) E0 sqrt(R/2), (4/3) Ecell sqrt(Rcell/2), (4/3) E0 sqrt(R/2) (2(R - Rcell))^1.5,
    in mpmath, rounded to double."""
    p0 = mp.mpf(4) / 3 * E0 * mp.sqrt(mp.mpf(R) / 2)
    p1 = mp.mpf(4) / 3 * ECELL * mp.sqrt(mp.mpf(RCELL) / 2)
    p2 = p0 * mp.power(2 * (mp.mpf(R) - mp.mpf(RCELL)), mp.mpf(1.5))
    return float(p0), float(p1), float(p2)


def _params(adhesion, ic_friction=False, prefactors=None):
    from bfm.core import Params, ADHESION_JKR_ISOTROPIC, ADHESION_NONE, JKR_EXACT
    p = Params()
    p.R, p.Rcell, p.E0 = R, RCELL, E0
    for k, v in enumerate(prefactors or _closed_form_prefactors()):
        p.repulsion_prefactors[k] = v
    p.nz_threshold = 1e-8
    p.cell_cell_coulomb_coeff = 1.0
    p.cell_surface_coulomb_coeff = 1.0
    p.no_surface = 1
    p.adhesion_mode = ADHESION_JKR_ISOTROPIC if adhesion else ADHESION_NONE
    p.jkr_mode = JKR_EXACT
    p.colidx_gamma = 21 if adhesion else -1
    p.friction_mode = 0
    p.ncols = 22
    p.colidx_eta_cell_cell = -1
    return p


# ----------------------------------------------------------------- energies
def _u_hertz_shell(delta):
    """Shell energy with the linear continuation past the shell thickness."""
    p0 = mp.mpf(4) / 3 * E0 * mp.sqrt(mp.mpf(R) / 2)
    dc = 2 * (mp.mpf(R) - mp.mpf(RCELL))
    if delta <= dc:
        return mp.mpf(2) / 5 * p0 * mp.power(delta, mp.mpf(2.5))
    return mp.mpf(2) / 5 * p0 * mp.power(dc, mp.mpf(2.5)) + p0 * mp.power(dc, mp.mpf(1.5)) * (delta - dc)


def _u_core(delta):
    dc = 2 * (mp.mpf(R) - mp.mpf(RCELL))
    if delta <= dc:
        return mp.mpf(0)
    return mp.mpf(8) / 15 * ECELL * mp.sqrt(mp.mpf(RCELL) / 2) * mp.power(delta - dc, mp.mpf(2.5))


def _jkr_radius(delta, gamma):
    """Larger real root of a^4 - 2 Req delta a^2 - (4 pi Req^2 gamma / E0) a + Req^2 delta^2."""
    req = mp.mpf(REQ)
    coeffs = [1, 0, -2 * req * delta, -4 * mp.pi * req ** 2 * gamma / E0, req ** 2 * delta ** 2]
    roots = mp.polyroots(coeffs, maxsteps=400, extraprec=200)
    real = [mp.re(z) for z in roots if abs(mp.im(z)) < mp.mpf(10) ** (-25)]
    return max(real)


def _u_jkr_at(delta, gamma):
    a = _jkr_radius(delta, gamma)
    req = mp.mpf(REQ)
    g = mp.mpf(gamma)
    return (mp.mpf(8) * E0 * a ** 5 / (15 * req ** 2)
            - mp.mpf(8) / 3 * mp.sqrt(mp.pi * g * E0) * mp.power(a, mp.mpf(3.5)) / req
            + 2 * mp.pi * g * a ** 2)


def _u_jkr(delta, gamma):
    dc = 2 * (mp.mpf(R) - mp.mpf(RCELL))
    if delta <= dc:
        return _u_jkr_at(delta, gamma)
    # linear continuation with the load at the cap
    a = _jkr_radius(dc, gamma)
    load = 4 * E0 * a ** 3 / (3 * mp.mpf(REQ)) - 4 * mp.sqrt(mp.pi * a ** 3 * gamma * E0)
    return _u_jkr_at(dc, gamma) + load * (delta - dc)


def _pair_energy(delta, gamma, adhesion, part="all"):
    if part == "core":
        return _u_core(delta)
    if not adhesion:
        return _u_hertz_shell(delta) + _u_core(delta)
    if gamma > 0:
        return _u_jkr(delta, mp.mpf(gamma)) + _u_core(delta)
    return _u_hertz_shell(delta) + _u_core(delta)


def _total_energy(rows, q, gammas, adhesion, part="all"):
    """rows: (i, j, d0 (3), s, t); q: dict cell -> 6 perturbations (dr, dn)."""
    tot = mp.mpf(0)
    for (i, j, d0, s, t) in rows:
        qi, qj = q[i], q[j]
        d = [mp.mpf(d0[c]) + (qj[c] + mp.mpf(t) * qj[3 + c]) - (qi[c] + mp.mpf(s) * qi[3 + c])
             for c in range(3)]
        dist = mp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        delta = 2 * mp.mpf(R) - dist
        if delta > 0:
            g = min(gammas[i], gammas[j])
            tot += _pair_energy(delta, g, adhesion, part)
    return tot


def _fd_gradient(rows, n, gammas, adhesion, part="all"):
    """n x 6 central-difference gradient of the total pair energy."""
    grad = np.zeros((n, 6))
    for cell in range(n):
        for c in range(6):
            def shifted(h):
                q = {k: [mp.mpf(0)] * 6 for k in range(n)}
                q[cell][c] = h
                return _total_energy(rows, q, gammas, adhesion, part)
            grad[cell, c] = float((shifted(H) - shifted(-H)) / (2 * H))
    return grad


# ------------------------------------------------------------------ geometry
def _unit(v):
    v = np.asarray(v, dtype=float)
    return v / np.linalg.norm(v)


def _cells(rs, ns, hls, gammas):
    n = len(rs)
    cells = np.zeros((n, 22))
    cells[:, 0] = np.arange(n)
    cells[:, 1:4] = rs
    cells[:, 4:7] = ns
    cells[:, 13] = 2 * np.asarray(hls)
    cells[:, 14] = hls
    cells[:, 16:21] = [1.5, 0.7, 720.0, 100.0, 1]
    cells[:, 21] = gammas
    return np.asfortranarray(cells)


def _place(orc, r1, n1, hl1, n2, hl2, offset, delta):
    """Second cell at r1 + offset, then moved along the distance vector so that
    the overlap is close to delta; returns r2 and the neighbor row (d, s, t)."""
    r2 = np.asarray(r1) + np.asarray(offset)
    for _ in range(4):
        d, s, t = orc.dist_between_cells(r1, n1, hl1, r2, n2, hl2)
        dist = np.linalg.norm(d)
        r2 = r2 + (2 * R - delta - dist) * d / dist
    d, s, t = orc.dist_between_cells(r1, n1, hl1, r2, n2, hl2)
    return r2, d, s, t


def _pair_case(orc, kind, delta, seed):
    rng = np.random.default_rng(seed)
    hl1, hl2 = rng.uniform(0.4, 1.6, 2)
    r1 = rng.uniform(-1, 1, 3) + np.array([0, 0, 5.0])
    n1 = _unit(rng.normal(size=3) * [1, 1, 1] + [0, 0, 0.3])
    if kind == "skew":
        n2 = _unit(n1 + rng.normal(scale=1.5, size=3))
        perp = _unit(np.cross(n1, n2))
        off = perp * 1.5 + rng.uniform(-0.3, 0.3) * n1
    else:  # near-parallel, |cos| >= 0.9999: the endpoint-averaging branch
        kick = _unit(np.cross(n1, rng.normal(size=3)))
        n2 = _unit(n1 + 0.005 * kick) * (1 if kind == "parallel" else -1)
        perp = _unit(np.cross(n1, _unit(rng.normal(size=3))))
        off = perp * 1.5 + rng.uniform(-0.4, 0.4) * n1
    r2, d, s, t = _place(orc, r1, n1, hl1, n2, hl2, off, delta)
    if kind != "skew":
        assert abs(np.dot(n1, n2)) >= 0.9999
    return (r1, n1, hl1), (r2, n2, hl2), (d, s, t)


# cases: (kind, overlap, seed). Overlaps inside the shell, past the shell
# (hard core) and close to the JKR-relevant range; none within 1e-6 of the cap.
CASES = [(k, dl, sd) for k in ("skew", "parallel", "antiparallel")
         for (dl, sd) in ((0.03, 1), (0.21, 2), (0.45, 3), (0.68, 4), (0.74, 5))]


def _row_matrix(rows):
    nb = np.zeros((len(rows), 7))
    for k, (i, j, d, s, t) in enumerate(rows):
        nb[k] = [i, j, d[0], d[1], d[2], s, t]
    return np.asfortranarray(nb)


def _rel(got, ref):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


def _check_pair(orc, kind, delta, seed, adhesion, gammas):
    c1, c2, (d, s, t) = _pair_case(orc, kind, delta, seed)
    cells = _cells([c1[0], c2[0]], [c1[1], c2[1]], [c1[2], c2[2]], gammas)
    rows = [(1, 0, -d, t, s)]     # reference order: i > j, d from i to j
    nb = _row_matrix(rows)
    got = orc.cell_cell_forces(cells, nb, _params(adhesion))
    ref = _fd_gradient(rows, 2, gammas, adhesion)
    return cells, nb, rows, got, ref


@pytest.mark.parametrize("kind,delta,seed", CASES)
def test_hertz_forces_are_energy_gradients(orc, kind, delta, seed):
    """Non-adhesive overload: shell Hertz + cap + hard core (mechanics.hpp:1209-1350)."""
    _, _, _, got, ref = _check_pair(orc, kind, delta, seed, False, [0.0, 0.0])
    assert _rel(got[:, :3], ref[:, :3]) < TOL, (got, ref)
    assert _rel(got[:, 3:], ref[:, 3:]) < TOL, (got, ref)
    assert np.allclose(got[0, :3], -got[1, :3], rtol=0, atol=1e-12 * np.abs(got).max())


@pytest.mark.parametrize("kind,delta,seed", CASES)
def test_jkr_forces_are_energy_gradients(orc, kind, delta, seed):
    """Adhesive overload, exact JKR radius (mechanics.hpp:1462-1523, adhesion.hpp:171-240),
    with min(gamma_i, gamma_j) (mechanics.hpp:1464) and the hard core (:1596-1645)."""
    gammas = [100.0, 250.0]
    _, _, _, got, ref = _check_pair(orc, kind, delta, seed, True, gammas)
    assert _rel(got[:, :3], ref[:, :3]) < TOL, (got, ref)
    assert _rel(got[:, 3:], ref[:, 3:]) < TOL, (got, ref)


@pytest.mark.parametrize("kind,delta,seed", CASES[::2])
def test_adhesive_overload_without_gamma_is_hertz(orc, kind, delta, seed):
    """gamma = 0 in the adhesive overload: Hertz with the EPS cap plus the separately
    accumulated hard core (mechanics.hpp:1526-1568, 1596-1645): the same energy."""
    _, _, _, got, ref = _check_pair(orc, kind, delta, seed, True, [0.0, 300.0])
    assert _rel(got, ref) < TOL


@pytest.mark.parametrize("delta", [0.02, 0.1, 0.3, 0.55, 0.7])
def test_jkr_tends_to_hertz_as_gamma_vanishes(orc, delta):
    """gamma -> 0+: the JKR load 4 E0 a^3/(3 Req) - 4 sqrt(pi a^3 gamma E0) with
    a -> sqrt(Req delta) is the Hertz force (4/3) E0 sqrt(R/2) delta^1.5."""
    c1, c2, (d, s, t) = _pair_case(orc, "skew", delta, 11)
    nb = _row_matrix([(1, 0, -d, t, s)])
    base = _cells([c1[0], c2[0]], [c1[1], c2[1]], [c1[2], c2[2]], [1e-12, 1e-12])
    jkr = orc.cell_cell_forces(base, nb, _params(True))
    hertz = orc.cell_cell_forces(base, nb, _params(False))
    assert _rel(jkr, hertz) < 1e-6


def test_three_cell_accumulation(orc):
    """Two rows that share a cell, once as i and once as j: per-cell sums of
    the rows' gradients land in the right rows."""
    rng = np.random.default_rng(5)
    r0, n0, h0 = np.array([0.0, 0.0, 4.0]), _unit([1, 0.2, 0.3]), 1.1
    n1, h1 = _unit([0.1, 1, 0.2]), 0.7
    r1, d01, s01, t01 = _place(orc, r0, n0, h0, n1, h1, [0.3, 1.5, 0.2], 0.25)
    n2, h2 = _unit([0.5, -0.4, 0.6]), 1.4
    r2, d12, s12, t12 = _place(orc, r1, n1, h1, n2, h2, [0.2, 1.4, -0.5], 0.12)
    gam = [80.0, 200.0, 30.0]
    cells = _cells([r0, r1, r2], [n0, n1, n2], [h0, h1, h2], gam)
    rows = [(1, 0, -d01, t01, s01), (2, 1, -d12, t12, s12)]
    nb = _row_matrix(rows)
    for adhesion in (False, True):
        got = orc.cell_cell_forces(cells, nb, _params(adhesion))
        ref = _fd_gradient(rows, 3, gam, adhesion)
        assert _rel(got, ref) < TOL, (adhesion, got, ref)
    del rng


@pytest.mark.parametrize("adhesion", [False, True])
def test_constrained_torques(orc, adhesion):
    """include_constraint = true (the velocity Verlet forms): Hertz, cap and
    hard-core torques are n x dE/dn ((s n) x F, mechanics.hpp:1298-1306,
    1608-1616); the JKR torque is the part of dE/dn perpendicular to n
    (s (-(n . F) n + F), adhesion.hpp:225-231). Both from the same FD gradient."""
    for kind, delta, seed in [("skew", 0.2, 21), ("skew", 0.7, 22), ("parallel", 0.4, 23)]:
        cells, nb, rows, _, ref = _check_pair(orc, kind, delta, seed, adhesion, [150.0, 150.0])
        got = orc.cell_cell_forces_ic(cells, nb, _params(adhesion))
        assert _rel(got[:, :3], ref[:, :3]) < TOL
        ns = [cells[0, 4:7], cells[1, 4:7]]
        if not adhesion:
            want = np.array([np.cross(ns[c], ref[c, 3:]) for c in range(2)])
            assert _rel(got[:, 3:], want) < TOL
        else:
            # the JKR part is projected, the hard-core part (past the cap) is crossed
            g_core = _fd_gradient(rows, 2, [150.0, 150.0], True, part="core")
            g_jkr = ref - g_core
            want = np.array([g_jkr[c, 3:] - np.dot(ns[c], g_jkr[c, 3:]) * ns[c]
                             + np.cross(ns[c], g_core[c, 3:]) for c in range(2)])
            assert _rel(got[:, 3:], want) < TOL
            assert (delta > DC) == bool(np.any(g_core != 0))


def test_repulsion_prefactors_closed_form():
    """bfm.core.repulsion_prefactors (simulation.hpp:1271-1274) against the
    closed forms in 40-digit arithmetic."""
    from bfm.core import repulsion_prefactors
    got = repulsion_prefactors(R, RCELL, E0, ECELL)
    want = _closed_form_prefactors()
    for g, w in zip(got, want):
        assert abs(g - w) <= 2e-16 * abs(w)
    # numerically: 3288.8... (shell), 1.2e6... (core), shell prefactor x 0.6^1.5
    assert abs(want[0] - 4 / 3 * 3900 * math.sqrt(0.4)) < 1e-9
    assert abs(want[2] / want[0] - 0.6 ** 1.5) < 1e-14


@pytest.mark.parametrize("mutant", ["prefactor", "torque_sign", "swap_st", "core_prefactor"])
def test_planted_mistakes_fail(orc, mutant):
    """Mistakes planted in what the oracle evaluates must fail the gradient check."""
    kind, delta, seed = ("skew", 0.7, 4)
    c1, c2, (d, s, t) = _pair_case(orc, kind, delta, seed)
    cells = _cells([c1[0], c2[0]], [c1[1], c2[1]], [c1[2], c2[2]], [0.0, 0.0])
    rows = [(1, 0, -d, t, s)]
    ref = _fd_gradient(rows, 2, [0.0, 0.0], False)
    pre = list(_closed_form_prefactors())
    nb_rows = rows
    if mutant == "prefactor":
        pre[0] *= 2
        pre[2] *= 2
    elif mutant == "core_prefactor":
        pre[1] *= 1.5
    elif mutant == "torque_sign":
        nb_rows = [(1, 0, -d, t, -s)] if abs(s) > abs(t) else [(1, 0, -d, -t, s)]
    elif mutant == "swap_st":
        nb_rows = [(1, 0, -d, s, t)]
    got = orc.cell_cell_forces(cells, _row_matrix(nb_rows), _params(False, prefactors=pre))
    assert _rel(got, ref) > 1e-3
