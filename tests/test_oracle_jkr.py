"""JKR contact radius and tables of the oracle.

Pinned by: mpmath 50-digit roots of the quartic (tests/golden/jkr_roots.json,
made by tests/golden/make_jkr_golden.py), and the reference's own checks:
Hertz limit a = sqrt(R delta) (test/src/testJKR.cpp:26-49), delta recovered from a
(:51-83), equilibrium a^3 = 9 pi gamma R^2 / E (:85-117), and the Parteli et al.
closed-form roots at delta=0.1, R=0.8, E=3900, gamma=100 (testContactArea.cpp:102-162).
"""
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _cases():
    with open(os.path.join(GOLD, "jkr_roots.json")) as f:
        return json.load(f)


def test_roots_match_mpmath_fixtures(orc):
    for c in _cases():
        a, it = orc.jkr_contact_radius(c["delta"], c["R"], c["E"], c["gamma"])
        if c["a"] is None:
            assert math.isnan(a)
        else:
            ref = float(c["a"])
            assert abs(a - ref) <= 4e-15 * ref, (c, a)
            assert it < 60


def test_host_table_solver_matches_fixtures():
    from bfm.jkr import contact_radius
    cs = [c for c in _cases() if c["a"] is not None and c["delta"] > 0]
    a = contact_radius(np.array([c["delta"] for c in cs]), np.array([c["R"] for c in cs]),
                       3900.0, np.array([c["gamma"] for c in cs]))
    ref = np.array([float(c["a"]) for c in cs])
    assert np.all(np.abs(a - ref) <= 4e-15 * ref)


def test_hertz_limit(orc):
    R = 0.8
    for d in np.linspace(0, 0.6, 100):
        a, _ = orc.jkr_contact_radius(d, R, 3900.0, 0.0)
        assert abs(a - math.sqrt(R * d)) < 1e-8


def test_overlap_recovered_and_above_hertz(orc):
    R, E = 0.8, 3900.0
    for g in np.linspace(10, 100, 41):
        for d in np.linspace(0, 0.6, 100):
            a, _ = orc.jkr_contact_radius(d, R, E, g)
            assert a >= math.sqrt(R * d)
            rec = a * a / R - 2 * math.sqrt(math.pi * g * a / E)
            assert abs(rec - d) < 1e-12


def test_equilibrium_radius(orc):
    """Zero force at a^3 = 9 pi gamma R^2 / E (testJKR.cpp:85-117)."""
    Req, E = 0.4, 3900.0
    for g in np.linspace(10, 100, 41):
        a_eq = (9 * math.pi * g * Req * Req / E) ** (1 / 3)
        d_eq = a_eq * a_eq / Req - 2 * math.sqrt(math.pi * g * a_eq / E)
        a, _ = orc.jkr_contact_radius(d_eq, Req, E, g)
        assert abs(a - a_eq) < 1e-12
        assert abs(orc.jkr_force_magnitude(a, Req, E, g)) < 1e-9 * (4 * E * a ** 3 / (3 * Req))


def test_parteli_closed_form(orc):
    delta, R, E, gamma = 0.1, 0.8, 3900.0, 100.0
    c0 = R * R * delta * delta
    c1 = -4 * math.pi * gamma * R * R / E
    c2 = -2 * R * delta
    P = -c2 * c2 / 12.0 - c0
    Q = -c2 ** 3 / 108.0 + c0 * c2 / 3.0 - c1 * c1 / 8.0
    U = (-Q / 2.0 + math.sqrt(Q * Q / 4.0 + P ** 3 / 27.0)) ** (1 / 3)
    s = -5 * c2 / 6.0 + (U - P / (3 * U))
    w = math.sqrt(c2 + 2 * s)
    lam = c1 / (2 * w)
    a1 = 0.5 * (w + math.sqrt(w * w - 4 * (c2 + s + lam)))
    a, _ = orc.jkr_contact_radius(delta, R, E, gamma)
    assert abs(a - a1) < 1e-13
    assert abs(a ** 4 + c2 * a * a + c1 * a + c0) < 1e-15


def test_linspace_matches_eigen_ends_and_host(orc):
    from bfm.jkr import linspaced
    for n in (2, 3, 10, 100, 101):
        v = orc.linspace(0.0, 0.6, n)
        assert v[0] == 0.0 and v[-1] == 0.6
        assert np.array_equal(v, linspaced(0.0, 0.6, n))
        assert np.all(np.diff(v) > 0)


def test_nearest_value_brute_force(orc):
    rng = np.random.default_rng(1)
    v = np.sort(rng.uniform(0, 1, 57))
    for x in np.concatenate([rng.uniform(-0.1, 1.1, 2000), v]):
        k = orc.nearest_value(v, x)
        best = np.min(np.abs(v - x))
        assert abs(v[k] - x) == best


def test_table_interpolation_error_bounded_by_mesh(orc):
    """Linear interpolation of a smooth radius curve: error O(h^2) of the 100-point mesh."""
    from bfm.jkr import reference_table
    R, Rcell, E, g = 0.8, 0.5, 3900.0, 100.0
    ov, rad = reference_table(R, Rcell, E, g)
    h = ov[1] - ov[0]
    for d in np.linspace(1e-3, 0.6, 997):
        r = orc.table_radius_1d(ov, rad, d)
        exact, _ = orc.jkr_contact_radius(d, R / 2, E, g)
        assert abs(r - exact) < 0.05 * h
    # at mesh points the table is exact
    for k in range(1, 99):
        assert abs(orc.table_radius_1d(ov, rad, ov[k]) - rad[k]) < 1e-15


def test_table_2d_bilinear(orc):
    from bfm.jkr import reference_table_2d
    R, Rcell, E = 0.8, 0.5, 3900.0
    ov, gm, rad = reference_table_2d(R, Rcell, E, 500.0, 40, 30)
    flat = np.ascontiguousarray(rad.ravel())
    # bilinear interpolation reproduces functions linear in each variable exactly
    lin = np.add.outer(2 * ov, 3e-4 * gm) + 0.1
    rng = np.random.default_rng(2)
    for d, g in zip(rng.uniform(0.01, 0.59, 300), rng.uniform(40, 499, 300)):  # a ~ gamma^(1/3) is steep near 0
        assert abs(orc.table_radius_2d(ov, gm, np.ascontiguousarray(lin.ravel()), d, g)
                   - (2 * d + 3e-4 * g + 0.1)) < 1e-12
        exact, _ = orc.jkr_contact_radius(d, R / 2, E, g)
        assert abs(orc.table_radius_2d(ov, gm, flat, d, g) - exact) < 5e-3
